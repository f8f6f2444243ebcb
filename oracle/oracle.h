/*
 * CPU oracle for the rt_entry query path: entry-point selection over kd-tree
 * leaf AABBs followed by greedy beam search on a proximity graph.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm, used by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg as the checker.  The product path (rt_entry_amd, librte)
 * never links, loads or calls it.
 *
 * Every function names the reference file:line it restates.  Parity
 * decisions where the reference is undefined or non-deterministic are listed
 * in orc_search() and in the repository's design notes.
 */
#ifndef RTE_ORACLE_H
#define RTE_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* (float)0x1fffffff: the reference's "empty slot" distance
 * (warpselect/structure_on_device.cuh:6, assigned to a float field). */
#define ORC_MAXD 536870912.0f

enum { ORC_MODE_REF = 0, ORC_MODE_TUNED = 1 };

typedef struct orc_params {
  int k;            /* results per query (topk), k <= beam */
  int beam;         /* n_candidates, a power of two (graph.cu:53) */
  int width;        /* search_width */
  int max_iter;     /* graph_search.cuh:976 */
  int mode;         /* ORC_MODE_REF: dist-only compares; TUNED: (dist,|id|) */
  int algo;         /* 1: leaf entry table; 0/2: per-thread LCG entries */
  int bloom_words;  /* 32-bit words of the visited-set filter */
  int bloom_hashes; /* hash functions of the visited-set filter */
} orc_params;

/* Bloom hash h (0..9) of id x (graph_search.cuh:13-22, constants :89-97),
 * before the modulo by the filter's bit count. */
uint32_t orc_hash_raw(int h, uint32_t x);

/* auto_tune_bloom.h:67-89 restated: filter sizing for a beam of nc. */
void orc_hash_parameter(int nc, int* hash_len, int* bit, int* hash);

/* kd-tree over 3-D points (kdtree.cpp:4-195).  Writes leaf boxes in
 * left-first DFS order (minx,miny,minz,maxx,maxy,maxz), leaf membership
 * belong[i] for every point, and the leaf count.  Returns 0, or -1 when a
 * split fails (kdtree.cpp:182-185), or -2 when more than max_leaves leaves
 * would be produced. */
int orc_kdtree_build(const float* pts3, int n, int max_node, float* aabbs,
                     int max_leaves, int* belong, int* n_leaf);

/* kdtree.cpp:103-125: the bisection median along one axis of a point set. */
float orc_find_xm(const float* vals, int n, float lo, float hi);

/* kdtree.cpp:91-101: per-leaf ascending point ids, padded with n. */
void orc_aabb_pids(const int* belong, int n, int n_leaf, int cap, int32_t* pids);

/* aabb.cu:27-38: grow each box by ratio * width on every axis. */
void orc_expand_aabbs(float* aabbs, int n_leaf, float ratio);

/* entry.cu:22: leaf capacity = (int)(np * point_ratio) in float32. */
int orc_leaf_capacity(int np, float point_ratio);

/* Entry rule replacing the OptiX ray cast (devicePrograms.cu:29-96):
 * hits[q] = smallest leaf id whose (expanded, closed) box contains
 * (pq[q*stride+0..2]); 0 if none (entry.cu:88-89). */
void orc_select_entry(const float* aabbs, int n_leaf, const float* pq, int64_t nq,
                      int stride, int32_t* hits);

/* FP64 (Q - mu) R[:, :d_out]: the closed form the projection GEMM
 * (graph.cu:262-263, matrix.cu:39-69) computes in FP32. */
void orc_project_f64(const float* Q, int64_t nq, int d, const float* R, int r_cols,
                     int d_out, const float* mu, double* out);

/* Squared L2 distance in the reference's association order
 * (graph_search.cuh:500-886): 32 lane partials over dims l+32j as an FMA
 * chain, then a shfl_down tree with offsets 16,8,4,2,1. */
float orc_distance(const float* p, const float* q, int d);

/* Beam search for nq queries (graph_search.cuh:56-1964).  graph rows have
 * stride 1<<shift, padded with np.  pids is the n_leaf x cap leaf table
 * (kdtree.cpp:91-101); hits the per-query leaf.  Outputs |id| and dist of
 * beam slots 0..k-1.  stats (may be NULL) receives per query
 * {distance evaluations, graph rows read, iterations}.  n_threads <= 0 uses
 * every OpenMP thread.  Returns 0 or a negative error. */
int orc_search(const float* base, int64_t np, int64_t stride, int d,
               const int32_t* graph, int shift, const float* queries, int64_t nq,
               int64_t q_stride, const int32_t* pids, int cap, const int32_t* hits,
               const orc_params* p, int32_t* ids_out, float* dists_out,
               int64_t* stats, int n_threads);

/* Per-query beam search with a beam, k and leaf of its own (the batched,
 * uneven form).  k_out_stride is the row stride of ids_out/dists_out. */
int orc_search_uneven(const float* base, int64_t np, int64_t stride, int d,
                      const int32_t* graph, int shift, const float* queries,
                      int64_t nq, int64_t q_stride, const int32_t* pids, int cap,
                      const int32_t* hits, const int32_t* k_per_q,
                      const int32_t* beam_per_q, const orc_params* p,
                      int k_out_stride, int32_t* ids_out, float* dists_out,
                      int n_threads);

/* Reference compare-exchange networks on (dist,id) arrays, exposed for the
 * network tests (graph_search.cuh:888-919 sort, :933-958 merge). */
void orc_sort_desc(float* d, int32_t* id, int m, int mode);
void orc_merge_asc(float* d, int32_t* id, int n, int mode);
/* One batch into the beam, as the search does it: sort B (m entries)
 * descending, min-exchange the min(n, m) pairs L[n-C+i] / B[m-C+i]
 * (skip_max: skip a batch entry at MAX, the main loop's rule), merge L
 * (n entries) ascending (graph_search.cuh:888-958, :1424-1493).  Both arrays
 * are updated in place. */
void orc_merge_batch(float* ld, int32_t* li, int n, float* bd, int32_t* bi, int m, int mode,
                     int skip_max);

/* Full-precision rerank (graph_helper.cu:146-213 Graph::parallel_reorder,
 * enabled by REORDER in graph.cu:326-336): distance of every candidate in
 * the original space as eight partial sums (dims j = l mod 8, sequential in
 * j, product then add, no fused multiply-add), the partials added in order
 * 0..7; candidates sorted by (dist, id) ascending (the reference's
 * std::sort leaves ties unspecified); the first k written.  Ids outside
 * [0, np) get dist +inf (the reference reads past the base).  d % 8 == 0
 * (the reference's remainder handling is broken otherwise). */
float orc_rerank_distance(const float* p, const float* q, int d);
int orc_rerank(const float* base, int64_t np, int64_t stride, int d, const float* queries,
               int64_t nq, int64_t q_stride, const int32_t* cand, int n_cand, int k,
               int32_t* ids_out, float* dists_out, int n_threads);

/* graph_helper.cu:64-121: recall@at of k result ids against gt rows. */
double orc_recall(const int32_t* ids, int64_t nq, int k, const int32_t* gt, int gt_k,
                  int at);

int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
