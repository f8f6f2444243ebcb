// Host kd-tree over the first three PCA coordinates, producing the leaf AABBs
// and the per-leaf point table that seed the search.
//
// Same leaf sets as the reference's Kdtree (kdtree.cpp:4-195):
//  - a node is a leaf when it holds at most max_node points (:171-174);
//  - otherwise it splits the widest axis of its tight box (:30-42, ties to the
//    lowest axis) at the bisection value of findxM (:103-125): halve [lo,hi]
//    until the width is below max((hi-lo)/1000, 1e-5) or exactly n/2 points
//    are <= mid, keeping the last mid with at most n/2 points at or below it;
//  - values below the split go left, the rest right (:44-62);
//  - leaf ids follow left-first DFS order (:189-194), each leaf box is the
//    tight box of its points (:127-138, 171-174);
//  - the point table lists each leaf's ids ascending, padded with np
//    (:91-101).
// The split is done with per-level counting passes parallelised with OpenMP;
// the order of points inside a node is irrelevant to every output.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "kdtree.h"

namespace rte {

namespace {

struct Node {
  int64_t l, r;
};

float find_split(const float* vals, int64_t n, float lo, float hi) {
  float l = lo, r = hi;
  float xm = (float)((double)(l + r) / 2.0);
  float eps = std::max((r - l) / 1000.0f, 0.00001f);
  const int64_t half = n / 2;
  while (r - l > eps) {
    const float mid = (float)((double)(r + l) / 2.0);
    int64_t cnt = 0;
#pragma omp parallel for reduction(+ : cnt) schedule(static) if (n > 65536)
    for (int64_t i = 0; i < n; ++i) cnt += vals[i] <= mid;
    if (cnt <= half) {
      l = mid;
      xm = mid;
      if (cnt == half) break;
    } else {
      r = mid;
    }
  }
  return xm;
}

}  // namespace

int build_kdtree(const float* pts3, int64_t np, int max_node, int max_leaves, float* aabbs,
                 int32_t* pids, int* n_leaf_out) {
  *n_leaf_out = 0;
  if (np <= 0 || max_node <= 0) return -1;
  std::vector<int32_t> idx(np), tmp(np);
  std::vector<float> vals(np);
  for (int64_t i = 0; i < np; ++i) idx[i] = (int32_t)i;
  std::vector<Node> stack;
  stack.push_back({0, np});
  int leaves = 0;
  while (!stack.empty()) {
    const Node t = stack.back();
    stack.pop_back();
    if (t.l >= t.r) continue;
    float bmin[3], bmax[3];
    for (int a = 0; a < 3; ++a) bmin[a] = bmax[a] = pts3[(int64_t)idx[t.l] * 3 + a];
    for (int64_t i = t.l + 1; i < t.r; ++i) {
      const float* p = pts3 + (int64_t)idx[i] * 3;
      for (int a = 0; a < 3; ++a) {
        bmin[a] = std::min(bmin[a], p[a]);
        bmax[a] = std::max(bmax[a], p[a]);
      }
    }
    const int64_t cnt = t.r - t.l;
    if (cnt <= max_node) {
      if (leaves >= max_leaves) return -2;
      float* box = aabbs + (int64_t)leaves * 6;
      for (int a = 0; a < 3; ++a) {
        box[a] = bmin[a];
        box[3 + a] = bmax[a];
      }
      std::sort(idx.begin() + t.l, idx.begin() + t.r);
      int32_t* row = pids + (int64_t)leaves * max_node;
      for (int64_t j = 0; j < max_node; ++j) row[j] = j < cnt ? idx[t.l + j] : (int32_t)np;
      ++leaves;
      continue;
    }
    int axis = 0;
    float maxw = 0.0f;
    for (int a = 0; a < 3; ++a)
      if (bmax[a] - bmin[a] > maxw) {
        maxw = bmax[a] - bmin[a];
        axis = a;
      }
    for (int64_t i = 0; i < cnt; ++i) vals[i] = pts3[(int64_t)idx[t.l + i] * 3 + axis];
    const float xm = find_split(vals.data(), cnt, bmin[axis], bmax[axis]);
    int64_t nl = 0, nr = 0;
    for (int64_t i = 0; i < cnt; ++i) {
      const int32_t id = idx[t.l + i];
      if (vals[i] < xm) idx[t.l + nl++] = id;
      else tmp[nr++] = id;
    }
    memcpy(idx.data() + t.l + nl, tmp.data(), sizeof(int32_t) * nr);
    const int64_t median = t.l + nl;
    if (median == t.l || median == t.r) return -3;  // kdtree.cpp:182-185 "split failed"
    stack.push_back({median, t.r});
    stack.push_back({t.l, median});
  }
  *n_leaf_out = leaves;
  return 0;
}

// aabb.cu:27-38
void expand_aabbs(float* aabbs, int n_leaf, float ratio) {
  for (int i = 0; i < n_leaf; ++i)
    for (int a = 0; a < 3; ++a) {
      float* mn = aabbs + (int64_t)i * 6 + a;
      float* mx = aabbs + (int64_t)i * 6 + 3 + a;
      const float w = *mx - *mn;
      const float grow = w * ratio;
      *mn = *mn - grow;
      *mx = *mx + grow;
    }
}

}  // namespace rte
