/*
 * CPU oracle for the rt_entry query path.  TEST INFRASTRUCTURE ONLY: used
 * by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker; the product library never links or calls it.
 *
 * Built with -ffp-contract=off so that the only fused multiply-adds are the
 * explicit fmaf() calls of orc_distance().
 *
 * Parity decisions (the reference is undefined or non-deterministic here):
 *  P1  visited-set filter zeroed at search start (graph_search.cuh:88 never
 *      clears it before first use);
 *  P2  first expansion mask = all ones over the first min(beam,32) slots
 *      (graph_search.cuh:970 shifts by 32 when beam >= 32; NVIDIA's shl
 *      clamps and yields all ones);
 *  P3  one row's neighbours are all tested against the filter before any of
 *      them is added (graph_search.cuh:998-1008 with lock-step lanes);
 *  P4  expanding a slot whose id is >= np (only reachable through P2 when the
 *      entry row has fewer than 32 valid ids) reads no graph row: the row is
 *      taken as all-np (the reference reads past the graph);
 *  P5  entry leaf = smallest leaf id whose expanded closed box contains the
 *      query's first three coordinates, 0 if none (OptiX any-hit order is
 *      unspecified, devicePrograms.cu:66-96, entry.cu:88-89);
 *  P6  compare-exchange networks run with a barrier between every substep
 *      (graph_search.cuh:892-918 has none inside a step; harmless while one
 *      warp does all pairs);
 *  P7  mode TUNED compares (dist, |id|) lexicographically; mode REF compares
 *      dist only with strict inequalities, as the reference.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* graph_search.cuh:89-97 */
static const uint32_t k_hash_consts[20] = {
    0x924ed183U, 0xd854fc0aU, 0xecf5e3b7U, 0x1bead407U, 0x28a30449U,
    0xbfc4d99fU, 0x715030e2U, 0xffcfb45bU, 0x6e4ce166U, 0xeb53c362U,
    0xa93c4f40U, 0xcecde0a4U, 0x0288592dU, 0x362c37bcU, 0x9d4824f0U,
    0xfdbdd68bU, 0x63258c85U, 0x6726905cU, 0x609500f9U, 0x4de48422U};

/* graph_search.cuh:13-22 (the mix), without the final modulo. */
uint32_t orc_hash_raw(int h, uint32_t x) {
  x ^= x >> 16;
  x *= k_hash_consts[h << 1];
  x ^= x >> 13;
  x *= k_hash_consts[(h << 1) + 1];
  x ^= x >> 16;
  return x;
}

/* ---- visited-set filter (graph_search.cuh:7-43) ---------------------- */
typedef struct {
  uint32_t* w;
  int words;
  int hashes;
  uint32_t bits;
} bloom_t;

static inline void bloom_add(bloom_t* b, uint32_t x) {
  for (int h = 0; h < b->hashes; ++h) {
    uint32_t pos = orc_hash_raw(h, x) % b->bits;          /* :21 */
    b->w[pos % (uint32_t)b->words] |= 1u << (pos / (uint32_t)b->words); /* :8-11 */
  }
}

static inline int bloom_test(const bloom_t* b, uint32_t x) {
  for (int h = 0; h < b->hashes; ++h) {
    uint32_t pos = orc_hash_raw(h, x) % b->bits;
    if (!((b->w[pos % (uint32_t)b->words] >> (pos / (uint32_t)b->words)) & 1u)) return 0; /* :25-28 */
  }
  return 1;
}

/* ---- auto_tune_bloom.h:8-89 ------------------------------------------ */
static double p_from_kr(double k, double r) { return pow(1 - exp(-k / r), k); }
static double k_from_r(double r) { return round(log(2) * r); }

void orc_hash_parameter(int nc, int* hash_len, int* bit, int* hash) {
  double n = (double)nc * 12.897;
  const int maxn = 11;
  for (int i = 1; i <= maxn; ++i) {
    double m = 32.0 * (double)(1 << i);
    double r = m / n;
    double k = k_from_r(r);
    double p = p_from_kr(k, r);
    /* auto_tune_bloom.h:83-87: the unbraced `if (i == maxn)` only guards the
     * first assignment, so every non-final pass also writes bit/hash; the
     * values returned are those of the pass that breaks (or the last). */
    if (p < 0.03) {
      *hash_len = 1 << i;
      *bit = i;
      *hash = (int)round(k);
      return;
    }
    if (i == maxn) *hash_len = 1 << i;
    *bit = i;
    *hash = (int)round(k);
  }
}

/* ---- kd-tree (kdtree.cpp) -------------------------------------------- */
typedef struct {
  int l, r;
} task_t;

/* kdtree.cpp:103-125.  vals holds the n coordinates of the node on the split
 * axis; lo/hi are the node's tight bounds on that axis. */
float orc_find_xm(const float* vals, int n, float lo, float hi) {
  float l = lo, r = hi;
  float xm = (float)((double)(l + r) / 2.0);
  float eps = (r - l) / 1000.0f;
  if (eps < 0.00001f) eps = 0.00001f;
  while (r - l > eps) {
    float mid = (float)((double)(r + l) / 2.0);
    int cnt = 0;
    for (int i = 0; i < n; ++i) cnt += vals[i] <= mid;
    if (cnt <= n / 2) {
      l = mid;
      xm = mid;
      if (cnt == n / 2) break;
    } else {
      r = mid;
    }
  }
  return xm;
}

int orc_kdtree_build(const float* pts3, int n, int max_node, float* aabbs,
                     int max_leaves, int* belong, int* n_leaf) {
  int* idx = (int*)malloc(sizeof(int) * (size_t)n);
  int* tmp = (int*)malloc(sizeof(int) * (size_t)n);
  float* vals = (float*)malloc(sizeof(float) * (size_t)n);
  /* the stack never holds more than one pending right sibling per level */
  task_t* stack = (task_t*)malloc(sizeof(task_t) * 256);
  int sp = 0, leaves = 0, rc = 0;
  for (int i = 0; i < n; ++i) idx[i] = i;
  if (n > 0) stack[sp++] = (task_t){0, n};
  while (sp > 0) {
    task_t t = stack[--sp];
    if (t.l >= t.r) continue;
    /* kdtree.cpp:127-138 tight box */
    float bmin[3], bmax[3];
    for (int a = 0; a < 3; ++a) bmin[a] = bmax[a] = pts3[(size_t)idx[t.l] * 3 + a];
    for (int i = t.l + 1; i < t.r; ++i)
      for (int a = 0; a < 3; ++a) {
        float v = pts3[(size_t)idx[i] * 3 + a];
        if (v < bmin[a]) bmin[a] = v; /* std::min keeps the first on ties */
        if (bmax[a] < v) bmax[a] = v;
      }
    if (t.r - t.l <= max_node) { /* kdtree.cpp:171-174, add_aabb :64-70 */
      if (leaves >= max_leaves) { rc = -2; break; }
      for (int a = 0; a < 3; ++a) {
        aabbs[(size_t)leaves * 6 + a] = bmin[a];
        aabbs[(size_t)leaves * 6 + 3 + a] = bmax[a];
      }
      for (int i = t.l; i < t.r; ++i) belong[idx[i]] = leaves;
      ++leaves;
      continue;
    }
    /* kdtree.cpp:30-42 widest axis, strict > keeps the lowest on ties */
    float maxw = 0.0f;
    int axis = 0;
    for (int a = 0; a < 3; ++a)
      if (bmax[a] - bmin[a] > maxw) { maxw = bmax[a] - bmin[a]; axis = a; }
    int cnt = t.r - t.l;
    for (int i = 0; i < cnt; ++i) vals[i] = pts3[(size_t)idx[t.l + i] * 3 + axis];
    float xm = orc_find_xm(vals, cnt, bmin[axis], bmax[axis]);
    /* kdtree.cpp:44-62: the two-pointer split leaves values < xm on the
     * left and values >= xm on the right; order inside each side does not
     * matter to anything downstream (boxes, counts and the ascending pid
     * table are order-free), so a stable partition is used here. */
    int nl = 0, nr = 0;
    for (int i = 0; i < cnt; ++i)
      if (vals[i] < xm) idx[t.l + nl++] = idx[t.l + i]; else tmp[nr++] = idx[t.l + i];
    memcpy(idx + t.l + nl, tmp, sizeof(int) * (size_t)nr);
    int median = t.l + nl;
    if (median == t.l || median == t.r) { rc = -1; break; } /* :182-185 */
    if (sp + 2 > 256) { rc = -3; break; }
    stack[sp++] = (task_t){median, t.r}; /* right pushed first: left-first DFS */
    stack[sp++] = (task_t){t.l, median};
  }
  *n_leaf = leaves;
  free(idx);
  free(tmp);
  free(vals);
  free(stack);
  return rc;
}

/* kdtree.cpp:91-101 */
void orc_aabb_pids(const int* belong, int n, int n_leaf, int cap, int32_t* pids) {
  int* count = (int*)calloc((size_t)n_leaf, sizeof(int));
  for (int64_t i = 0; i < (int64_t)n_leaf * cap; ++i) pids[i] = n;
  for (int i = 0; i < n; ++i) {
    int a = belong[i];
    if (count[a] < cap) pids[(int64_t)a * cap + count[a]] = i;
    count[a]++;
  }
  free(count);
}

/* aabb.cu:27-38 */
void orc_expand_aabbs(float* aabbs, int n_leaf, float ratio) {
  for (int i = 0; i < n_leaf; ++i)
    for (int a = 0; a < 3; ++a) {
      float* mn = &aabbs[(size_t)i * 6 + a];
      float* mx = &aabbs[(size_t)i * 6 + 3 + a];
      float w = *mx - *mn;
      *mn -= w * ratio;
      *mx += w * ratio;
    }
}

/* entry.cu:22 (float point_ratio, int np) */
int orc_leaf_capacity(int np, float point_ratio) {
  float c = (float)np * point_ratio;
  return (int)c;
}

/* P5; devicePrograms.cu:33-35 reads queries[q*dim + 0..2]. */
void orc_select_entry(const float* aabbs, int n_leaf, const float* pq, int64_t nq,
                      int stride, int32_t* hits) {
#pragma omp parallel for schedule(static)
  for (int64_t q = 0; q < nq; ++q) {
    const float* x = pq + q * stride;
    int hit = 0;
    for (int i = 0; i < n_leaf; ++i) {
      const float* b = aabbs + (size_t)i * 6;
      if (x[0] >= b[0] && x[0] <= b[3] && x[1] >= b[1] && x[1] <= b[4] && x[2] >= b[2] &&
          x[2] <= b[5]) {
        hit = i;
        break;
      }
    }
    hits[q] = hit;
  }
}

void orc_project_f64(const float* Q, int64_t nq, int d, const float* R, int r_cols,
                     int d_out, const float* mu, double* out) {
#pragma omp parallel for schedule(static)
  for (int64_t q = 0; q < nq; ++q)
    for (int c = 0; c < d_out; ++c) {
      double s = 0.0;
      for (int j = 0; j < d; ++j)
        s += ((double)Q[q * d + j] - (double)mu[j]) * (double)R[(size_t)j * r_cols + c];
      out[q * d_out + c] = s;
    }
}

/* graph_search.cuh:686-878 (delta, FMA-contracted sum) and :873-878 (tree) */
float orc_distance(const float* p, const float* q, int d) {
  float acc[32];
  for (int l = 0; l < 32; ++l) {
    float a = 0.0f;
    for (int j = l; j < d; j += 32) {
      float t = p[j] - q[j];
      a = fmaf(t, t, a);
    }
    acc[l] = a;
  }
  for (int o = 16; o >= 1; o >>= 1)
    for (int l = 0; l < o; ++l) acc[l] = acc[l] + acc[l + o];
  return acc[0];
}

/* ---- compare-exchange networks --------------------------------------- */
typedef struct {
  float d;
  int32_t id;
} pair_t;

static inline int key_gt(pair_t a, pair_t b, int mode) {
  if (mode == ORC_MODE_REF) return a.d > b.d;
  if (a.d != b.d) return a.d > b.d;
  return abs(a.id) > abs(b.id);
}

static inline void cswap(pair_t* a, pair_t* b) {
  pair_t t = *a;
  *a = *b;
  *b = t;
}

/* graph_search.cuh:888-919: bitonic sort, descending overall. */
static void sort_desc(pair_t* a, int m, int mode) {
  for (int step = 1; step <= m / 2; step *= 2)
    for (int sub = step; sub >= 1; sub /= 2)
      for (int i = 0; i < m / 2; ++i) {
        int u = (i / sub) * 2 * sub + (i & (sub - 1));
        if ((i / step) % 2 == 0) {
          if (key_gt(a[u + sub], a[u], mode)) cswap(&a[u], &a[u + sub]);
        } else {
          if (key_gt(a[u], a[u + sub], mode)) cswap(&a[u], &a[u + sub]);
        }
      }
}

/* graph_search.cuh:933-958: half-cleaner cascade with step n/2 (every pair
 * ascending). */
static void merge_asc(pair_t* a, int n, int mode) {
  for (int sub = n / 2; sub >= 1; sub /= 2)
    for (int i = 0; i < n / 2; ++i) {
      int u = (i / sub) * 2 * sub + (i & (sub - 1));
      if (key_gt(a[u], a[u + sub], mode)) cswap(&a[u], &a[u + sub]);
    }
}

void orc_sort_desc(float* d, int32_t* id, int m, int mode) {
  pair_t* a = (pair_t*)malloc(sizeof(pair_t) * (size_t)m);
  for (int i = 0; i < m; ++i) a[i] = (pair_t){d[i], id[i]};
  sort_desc(a, m, mode);
  for (int i = 0; i < m; ++i) { d[i] = a[i].d; id[i] = a[i].id; }
  free(a);
}

void orc_merge_asc(float* d, int32_t* id, int n, int mode) {
  pair_t* a = (pair_t*)malloc(sizeof(pair_t) * (size_t)n);
  for (int i = 0; i < n; ++i) a[i] = (pair_t){d[i], id[i]};
  merge_asc(a, n, mode);
  for (int i = 0; i < n; ++i) { d[i] = a[i].d; id[i] = a[i].id; }
  free(a);
}

/* One batch into the beam: sort B descending (:888-919 / :1424-1453), then
 * min-exchange the C = min(N, M) pairs L[N-C+i] / B[M-C+i] (:921-931 /
 * :1455-1467; in the main loop a MAX batch element is skipped), then the
 * half-cleaner merge of L (:933-958 / :1469-1493).  When C = N the pairs
 * join L ascending with the N smallest of B descending, a bitonic sequence
 * whose merge leaves L sorted with the N smallest of L and B. */
static void merge_batch(pair_t* L, int N, pair_t* B, int M, int mode, int skip_max) {
  const int C = N < M ? N : M; /* :80-83 n_compare */
  sort_desc(B, M, mode);
  for (int i = 0; i < C; ++i) {
    int u = N - C + i;
    if (skip_max && B[M - C + i].d == ORC_MAXD) continue;
    if (key_gt(L[u], B[M - C + i], mode)) cswap(&L[u], &B[M - C + i]);
  }
  merge_asc(L, N, mode);
}

void orc_merge_batch(float* ld, int32_t* li, int N, float* bd, int32_t* bi, int M, int mode,
                     int skip_max) {
  pair_t* L = (pair_t*)malloc(sizeof(pair_t) * (size_t)(N + M));
  pair_t* B = L + N;
  for (int i = 0; i < N; ++i) L[i] = (pair_t){ld[i], li[i]};
  for (int i = 0; i < M; ++i) B[i] = (pair_t){bd[i], bi[i]};
  merge_batch(L, N, B, M, mode, skip_max);
  for (int i = 0; i < N; ++i) { ld[i] = L[i].d; li[i] = L[i].id; }
  for (int i = 0; i < M; ++i) { bd[i] = B[i].d; bi[i] = B[i].id; }
  free(L);
}

/* ---- the search (graph_search.cuh:56-1964) --------------------------- */
typedef struct {
  const float* base;
  int64_t np, stride;
  int d;
  const int32_t* graph;
  int shift;
  const int32_t* pids;
  int cap;
} space_t;

/* graph_search.cuh:1014-1031 / :1495-1516: flags of one 32-slot chunk with
 * the per-lane check_zero side effect (point 0 cannot carry the negative
 * "expanded" mark). */
static uint32_t scan_chunk(const pair_t* L, int n, int chunk, int64_t np,
                           uint32_t* czero) {
  uint32_t m = 0;
  for (int l = 0; l < 32; ++l) {
    int s = chunk * 32 + l;
    if (s >= n) continue;
    int32_t id = L[s].id;
    if (id > 0 && id < np) m |= 1u << l;
    else if (id == 0 && !(*czero & (1u << l))) {
      *czero |= 1u << l;
      m |= 1u << l;
    }
  }
  return m;
}

static void search_one(const space_t* S, const float* q, int64_t qidx, int hit, int N, int k,
                       const orc_params* p, pair_t* L, pair_t* B, uint32_t* bw,
                       int32_t* ids_out, float* dists_out, int64_t* st) {
  const int64_t np = S->np;
  const int D = 1 << S->shift;
  const int W = p->width;
  const int M = W * D;
  const int mode = p->mode;
  bloom_t bl = {bw, p->bloom_words, p->bloom_hashes, (uint32_t)p->bloom_words * 32u};
  int64_t E = 0, X = 0;
  memset(bw, 0, sizeof(uint32_t) * (size_t)p->bloom_words); /* P1 */
  for (int i = 0; i < N; ++i) L[i] = (pair_t){ORC_MAXD, (int32_t)np}; /* :473-477 */
  for (int i = 0; i < M; ++i) B[i] = (pair_t){ORC_MAXD, (int32_t)np};
  uint32_t seed[64];
  for (int t = 0; t < 64; ++t) seed[t] = (uint32_t)(qidx * 64 + t); /* :66 */

  /* entry seeding, :479-959 */
  int nch = (N + M - 1) / M;
  for (int c = 0; c < nch; ++c) {
    for (int i = 0; i < M; ++i) {
      int idx = c * M + i;
      if (idx < N) {
        int32_t id;
        if (p->algo == 1) {
          id = idx < S->cap ? S->pids[(int64_t)hit * S->cap + idx] : (int32_t)np;
        } else { /* :50-54, one LCG per thread t = i % 64 */
          uint32_t* sd = &seed[i & 63];
          *sd = *sd * 1103515245u + 12345u;
          id = (int32_t)(*sd % (uint32_t)np);
        }
        B[i].id = id;
        bloom_add(&bl, (uint32_t)id);
      } else {
        B[i].id = (int32_t)np;
      }
      if (B[i].id >= np) B[i].d = ORC_MAXD;
    }
    for (int i = 0; i < M; ++i)
      if (B[i].id < np) {
        B[i].d = orc_distance(S->base + (int64_t)B[i].id * S->stride, q, S->d);
        ++E;
      }
    merge_batch(L, N, B, M, mode, 0); /* no MAX skip while seeding */
  }

  /* main loop, :964-1532 */
  int iter = 0, done = 0, chunk = 0, hash_it = 0;
  uint32_t czero = 0;
  uint32_t mask = N >= 32 ? 0xffffffffu : ((1u << N) - 1u); /* P2 */
  const int nchunks = (N + 31) / 32;
  int32_t row[64 * 16];
  int pass[64 * 16];
  while (!done && iter < p->max_iter) {
    ++iter;
    for (int i = 0; i < M; ++i) B[i] = (pair_t){ORC_MAXD, (int32_t)np};
    int cnt = 0;
    while (cnt < W && mask != 0) {
      int b = __builtin_ctz(mask);
      int s = chunk * 32 + b;
      int32_t v = abs(L[s].id);
      L[s].id = -L[s].id;
      mask &= ~(1u << b);
      if (v < np) { /* P4 */
        memcpy(row, S->graph + ((int64_t)v << S->shift), sizeof(int32_t) * (size_t)D);
        ++X;
      } else {
        for (int i = 0; i < D; ++i) row[i] = (int32_t)np;
      }
      for (int i = 0; i < D; ++i) /* P3: test the whole row first */
        pass[i] = !(row[i] == np || bloom_test(&bl, (uint32_t)row[i]));
      for (int i = 0; i < D; ++i) {
        if (pass[i]) {
          B[cnt * D + i].id = row[i];
          bloom_add(&bl, (uint32_t)row[i]);
        } else {
          B[cnt * D + i].id = (int32_t)np;
        }
      }
      ++cnt;
      if (cnt == W) break;
      while (mask == 0 && chunk < nchunks) { /* :1014-1031 */
        ++chunk;
        mask = scan_chunk(L, N, chunk, np, &czero);
      }
    }
    if (cnt == 0) break;
    for (int i = 0; i < cnt * D; ++i) { /* :1036-1422 */
      if (B[i].id >= np) {
        B[i].d = ORC_MAXD;
      } else {
        B[i].d = orc_distance(S->base + (int64_t)B[i].id * S->stride, q, S->d);
        ++E;
      }
    }
    merge_batch(L, N, B, M, mode, 1);
    for (chunk = 0; chunk < nchunks; ++chunk) { /* :1495-1516 */
      mask = scan_chunk(L, N, chunk, np, &czero);
      if (mask != 0) break;
      if (chunk == nchunks - 1) done = 1;
    }
    if (hash_it == 4) { /* :1518-1531 */
      memset(bw, 0, sizeof(uint32_t) * (size_t)p->bloom_words);
      for (int j = 0; j < N; ++j) {
        int32_t x = abs(L[j].id);
        if (x < np) bloom_add(&bl, (uint32_t)x);
      }
      hash_it = 0;
    }
    ++hash_it;
  }
  for (int j = 0; j < k; ++j) { /* :1955-1963, plus distances */
    ids_out[j] = abs(L[j].id);
    if (dists_out) dists_out[j] = L[j].d;
  }
  if (st) {
    st[0] = E;
    st[1] = X;
    st[2] = iter;
  }
}

static int check_common(int64_t np, int d, int shift, const orc_params* p, int auto_bloom) {
  if (np <= 0 || np >= 0x7fffffff || d <= 0 || d > 960) return -1;
  if (shift < 0 || shift > 6) return -2;
  if (p->width < 1 || (p->width << shift) > 1024) return -3;
  if (!(auto_bloom && p->bloom_words == 0) &&
      (p->bloom_words <= 0 || p->bloom_words > 2048 || (p->bloom_words & (p->bloom_words - 1)) ||
       p->bloom_hashes < 1 || p->bloom_hashes > 10))
    return -4;
  if (p->max_iter < 0) return -5;
  return 0;
}

static int is_pow2(int x) { return x > 0 && (x & (x - 1)) == 0; }

int orc_search(const float* base, int64_t np, int64_t stride, int d,
               const int32_t* graph, int shift, const float* queries, int64_t nq,
               int64_t q_stride, const int32_t* pids, int cap, const int32_t* hits,
               const orc_params* p, int32_t* ids_out, float* dists_out,
               int64_t* stats, int n_threads) {
  int rc = check_common(np, d, shift, p, 0);
  if (rc) return rc;
  if (!is_pow2(p->beam) || p->beam > 1024 || p->k < 1 || p->k > p->beam) return -6;
  space_t S = {base, np, stride, d, graph, shift, pids, cap};
  const int N = p->beam, M = p->width << shift;
#ifdef _OPENMP
  if (n_threads <= 0) n_threads = omp_get_max_threads();
#else
  (void)n_threads;
#endif
#pragma omp parallel num_threads(n_threads)
  {
    pair_t* L = (pair_t*)malloc(sizeof(pair_t) * (size_t)N);
    pair_t* B = (pair_t*)malloc(sizeof(pair_t) * (size_t)M);
    uint32_t* bw = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)p->bloom_words);
#pragma omp for schedule(dynamic, 16)
    for (int64_t q = 0; q < nq; ++q) {
      int hit = (p->algo == 1) ? hits[q] : 0;
      search_one(&S, queries + q * q_stride, q, hit, N, p->k, p, L, B, bw,
                 ids_out + q * p->k, dists_out ? dists_out + q * p->k : NULL,
                 stats ? stats + q * 3 : NULL);
    }
    free(L);
    free(B);
    free(bw);
  }
  return 0;
}

int orc_search_uneven(const float* base, int64_t np, int64_t stride, int d,
                      const int32_t* graph, int shift, const float* queries,
                      int64_t nq, int64_t q_stride, const int32_t* pids, int cap,
                      const int32_t* hits, const int32_t* k_per_q,
                      const int32_t* beam_per_q, const orc_params* p,
                      int k_out_stride, int32_t* ids_out, float* dists_out,
                      int n_threads) {
  int rc = check_common(np, d, shift, p, p->mode == ORC_MODE_TUNED);
  if (rc) return rc;
  space_t S = {base, np, stride, d, graph, shift, pids, cap};
  const int M = p->width << shift;
  for (int64_t q = 0; q < nq; ++q)
    if (!is_pow2(beam_per_q[q]) || beam_per_q[q] > 1024 || k_per_q[q] < 1 ||
        k_per_q[q] > beam_per_q[q] || k_per_q[q] > k_out_stride)
      return -6;
#ifdef _OPENMP
  if (n_threads <= 0) n_threads = omp_get_max_threads();
#else
  (void)n_threads;
#endif
#pragma omp parallel num_threads(n_threads)
  {
    pair_t* L = (pair_t*)malloc(sizeof(pair_t) * 1024);
    pair_t* B = (pair_t*)malloc(sizeof(pair_t) * (size_t)M);
    uint32_t* bw = (uint32_t*)malloc(sizeof(uint32_t) * 2048);
#pragma omp for schedule(dynamic, 16)
    for (int64_t q = 0; q < nq; ++q) {
      orc_params pq = *p;
      pq.beam = beam_per_q[q];
      pq.k = k_per_q[q];
      if (p->mode == ORC_MODE_TUNED && p->bloom_words == 0) {
        int hl, bit, hh;
        orc_hash_parameter(pq.beam, &hl, &bit, &hh);
        pq.bloom_words = hl;
        pq.bloom_hashes = hh;
      }
      int hit = (p->algo == 1) ? hits[q] : 0;
      search_one(&S, queries + q * q_stride, q, hit, pq.beam, pq.k, &pq, L, B, bw,
                 ids_out + q * k_out_stride,
                 dists_out ? dists_out + q * k_out_stride : NULL, NULL);
    }
    free(L);
    free(B);
    free(bw);
  }
  return 0;
}

/* graph_helper.cu:161-190: eight accumulators over dims l, l + 8, ...
 * (mul then add), then accumulators 0..7 summed in order. */
float orc_rerank_distance(const float* p, const float* q, int d) {
  float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int j = 0; j < d; j += 8)
    for (int l = 0; l < 8; ++l) {
      const float diff = p[j + l] - q[j + l];
      const float sq = diff * diff;
      acc[l] = acc[l] + sq;
    }
  float dis = 0.0f;
  for (int l = 0; l < 8; ++l) dis = dis + acc[l];
  return dis;
}

typedef struct {
  float d;
  int32_t id;
} orc_pair;

static int pair_less(const orc_pair* a, const orc_pair* b) {
  return a->d < b->d || (a->d == b->d && a->id < b->id);
}

int orc_rerank(const float* base, int64_t np, int64_t stride, int d, const float* queries,
               int64_t nq, int64_t q_stride, const int32_t* cand, int n_cand, int k,
               int32_t* ids_out, float* dists_out, int n_threads) {
  if (d % 8 != 0 || d <= 0 || k < 1 || k > n_cand || stride < d || q_stride < d) return -1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(n_threads > 0 ? n_threads : 1)
  for (int64_t q = 0; q < nq; ++q) {
    orc_pair* v = (orc_pair*)malloc(sizeof(orc_pair) * (size_t)n_cand);
    const float* qv = queries + q * q_stride;
    for (int j = 0; j < n_cand; ++j) {
      const int32_t id = cand[q * n_cand + j];
      v[j].id = id;
      v[j].d = (id >= 0 && id < np) ? orc_rerank_distance(base + (int64_t)id * stride, qv, d)
                                    : INFINITY;
    }
    /* insertion sort by (dist, id): n_cand <= 512 */
    for (int i = 1; i < n_cand; ++i) {
      orc_pair x = v[i];
      int j = i - 1;
      while (j >= 0 && pair_less(&x, &v[j])) {
        v[j + 1] = v[j];
        --j;
      }
      v[j + 1] = x;
    }
    for (int j = 0; j < k; ++j) {
      ids_out[q * k + j] = v[j].id;
      if (dists_out) dists_out[q * k + j] = v[j].d;
    }
    free(v);
  }
  return 0;
}

/* graph_helper.cu:64-121 */
double orc_recall(const int32_t* ids, int64_t nq, int k, const int32_t* gt, int gt_k,
                  int at) {
  int64_t hits = 0;
  for (int64_t q = 0; q < nq; ++q)
    for (int i = 0; i < gt_k && i < at; ++i) {
      int32_t g = gt[q * gt_k + i];
      for (int j = 0; j < k; ++j)
        if (ids[q * k + j] == g) { ++hits; break; }
    }
  return (double)hits / (double)nq / (double)at;
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
