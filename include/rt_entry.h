/*
 * rt_entry for MI355X (gfx950): C ABI of the query path
 *   PCA query projection -> entry-leaf selection -> greedy beam search.
 *
 * Conventions
 *  - Every call returns RTE_OK (0) or a negative RTE_E* code; the message of
 *    the last failure on a context is rte_last_error(ctx).  Nothing calls
 *    exit() (the reference prints and exits, helper.h:4-34).
 *  - One context owns one HIP stream and the device buffers it allocated.
 *    The stream is a blocking stream (ordered with the null stream); a
 *    caller stream set with rte_set_stream() is ordered only by the caller.
 *    Calls on one context are not thread-safe; separate contexts are.
 *  - `where` says whether a pointer argument is host (RTE_HOST) or device
 *    (RTE_DEVICE) memory.  Host inputs are copied into context-owned device
 *    memory.  Device inputs to the rte_set_* calls are borrowed, not copied:
 *    the caller keeps them alive and unchanged until the context is destroyed
 *    or the slot is set again.  Host outputs are written before the call
 *    returns; device outputs are written in stream order on the context's
 *    stream (rte_synchronize() waits for them).
 *  - Ids are int32, distances are squared L2 in FP32.
 *  - A batch of zero queries (nq == 0) returns RTE_OK and touches nothing;
 *    its data pointers may be NULL (an empty device tensor has none).
 *    Scalar arguments such as strides are still checked.
 */
#ifndef RT_ENTRY_H
#define RT_ENTRY_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RTE_OK 0
#define RTE_EINVAL (-1)   /* bad argument or shape */
#define RTE_ESTATE (-2)   /* a required rte_set_* call is missing */
#define RTE_EHIP (-3)     /* HIP runtime error */
#define RTE_ENOMEM (-4)   /* device allocation failed */
#define RTE_EUNSUP (-5)   /* shape not compiled into this build */
#define RTE_EBUILD (-6)   /* host-side build (kd-tree split) failed */

#define RTE_HOST 0
#define RTE_DEVICE 1

#define RTE_MODE_REF 0   /* reference semantics: dist-only strict compares */
#define RTE_MODE_TUNED 1 /* (dist, id) compares, auto-sized visited filter */

typedef struct rte_ctx rte_ctx;

/* Search parameters (graph.cu:32-64 Graph ctor, flags.cpp defaults). */
typedef struct rte_params {
  int k;            /* results per query, 1 <= k <= beam (topk) */
  int beam;         /* n_candidates; rounded up to a power of two, 1..512
                       (the batched calls use at least 32) */
  int width;        /* search_width: graph rows expanded per iteration */
  int max_iter;     /* iteration cap (flags.cpp max_iter, default 100) */
  int mode;         /* RTE_MODE_REF or RTE_MODE_TUNED */
  int algo;         /* 1: leaf entry table; 0 or 2: per-lane LCG entries */
  int bloom_words;  /* 0: default (REF 512, TUNED auto_tune_bloom sizing) */
  int bloom_hashes; /* 0, or the mode's count: REF 3, TUNED 7 */
  int wg_waves;     /* queries (wave64s) per workgroup, 1..4; 0: 4 (with
                       wpq > 1 the workgroup is the wpq waves of one query) */
  int grid;         /* workgroups launched, > 0; or RTE_GRID_* (0: one per
                       query unit for rte_search, the resident maximum for
                       the batched calls) */
  int wpq;          /* wavefronts per query: 0 or 1 (one wave holds a query),
                       2 or 4 (the waves split each batch's distances and
                       meet at a workgroup barrier; for small batches) */
  int order;        /* batched calls: RTE_ORDER_GIVEN (0) or
                       RTE_ORDER_LEAF: queries sorted by entry leaf on the
                       device, the sorted range cut into one piece per XCD
                       (workgroups blockIdx % 8), so queries that gather
                       the same rows share an L2; never changes results */
  int queue;        /* where the beam (candidate queue) lives: RTE_QUEUE_VGPR
                       (0, registers) or RTE_QUEUE_LDS (an LDS queue of
                       64 * ceil(beam bound / 64) slots per wave, 8 bytes a
                       slot; fewer VGPRs for large beams, wpq 0 or 1 only);
                       never changes results */
} rte_params;

/* rte_params.grid codes, resolved per call from the launched kernel's
 * occupancy (resident workgroups per CU x CUs). */
#define RTE_GRID_DEFAULT 0    /* rte_search: one workgroup per query unit;
                                 batched: the resident maximum */
#define RTE_GRID_RESIDENT2 (-1) /* twice the resident maximum */
#define RTE_GRID_BALANCED (-2)  /* the fewest units that take the same number
                                   of rounds through the queue as the resident
                                   maximum */
#define RTE_GRID_PER_QUERY (-3) /* one workgroup per query unit */

#define RTE_ORDER_GIVEN 0
#define RTE_ORDER_LEAF 1

#define RTE_QUEUE_VGPR 0
#define RTE_QUEUE_LDS 1

/* ---- context ---------------------------------------------------------- */
/* Replaces Graph::Graph + cublasCreate (graph.cu:32-64). */
int rte_create(rte_ctx** ctx, int device);
/* Replaces Graph::CleanUp / OptiXRT::CleanUp (graph.cu:339-342,
 * raytracing.cpp:263-281). */
int rte_destroy(rte_ctx* ctx);
const char* rte_last_error(const rte_ctx* ctx);
/* Use the caller's HIP stream (hipStream_t) instead of the context's own;
 * NULL restores the context stream. */
int rte_set_stream(rte_ctx* ctx, void* stream);
void* rte_get_stream(rte_ctx* ctx);
int rte_synchronize(rte_ctx* ctx);

/* ---- data ------------------------------------------------------------- */
/* Search-space vectors, np x d row-major with row stride `stride` floats
 * (PCA-rotated or raw).  Replaces Graph::Input / Projection's upload of
 * d_points_ / d_pca_points (graph.cu:73-112, 180-186). */
int rte_set_base(rte_ctx* ctx, const float* base, int64_t np, int d, int64_t stride,
                 int where);
/* Proximity graph rows, np x degree int32, padded with np.  Rows are
 * re-laid out to a power-of-two stride (graph.cu:134 offset_shift) when
 * degree is not one.  Replaces Graph::RB_Graph (graph.cu:114-139). */
int rte_set_graph(rte_ctx* ctx, const int32_t* rows, int64_t np, int degree, int where);
/* PCA rotation R (d x r_cols row-major; column j = eigenvector j) and mean
 * mu (d); the projection keeps the first d_out columns.  Replaces
 * Graph::Projection's device setup (graph.cu:188-228: copy of R[:, :DIM],
 * mu.R GEMM and replicateVector). */
int rte_set_pca(rte_ctx* ctx, const float* R, const float* mu, int d, int r_cols,
                int d_out, int where);
/* PQ = Q R[:, :d_out] - mu R[:, :d_out] as an FP32 MFMA GEMM.  Replaces
 * matrixMultiply(handle_, d_queries_, d_rotation, d_pca_queries, nq, DIM,
 * dim_, 1, -1) (graph.cu:262-263, matrix.cu:39-69). */
int rte_project(rte_ctx* ctx, const float* Q, int64_t nq, float* PQ, int where);

/* ---- entry selection --------------------------------------------------- */
/* Leaf boxes (n_leaf x {minx,miny,minz,maxx,maxy,maxz}, already expanded)
 * and the leaf point table (n_leaf x cap ascending ids, padded with np).
 * Builds the traversal hierarchy over the boxes.  Replaces
 * RT_Entry::InitRT -> OptiXRT::Setup/BuildAccel (entry.cu:79-98,
 * raytracing.cpp:184-261) and the aabb_entries upload (entry.cu:60-69). */
int rte_set_leaves(rte_ctx* ctx, const float* aabbs, const int32_t* pids, int n_leaf,
                   int cap, int where);
/* Host kd-tree over 3-D points (np x 3): leaves of at most max_node points
 * in left-first DFS order; writes unexpanded boxes (n_leaf x 6, capacity
 * max_leaves) and the ascending pid table (max_leaves x max_node).
 * Replaces Kdtree + computeAabbPid (kdtree.cpp:4-195) in
 * RT_Entry::BlockUp (entry.cu:8-77). */
int rte_build_kdtree(const float* pts3, int64_t np, int max_node, int max_leaves,
                     float* aabbs_out, int32_t* pids_out, int* n_leaf_out);
/* The same kd-tree built on the GPU: identical leaf sets, boxes (as values),
 * DFS leaf order, pid table and error codes as rte_build_kdtree.  pts holds
 * np rows of `stride` floats whose first three are the coordinates (e.g. the
 * PCA base itself); pts, aabbs_out and pids_out are host or device per
 * `where`.  max_node <= 2048.  Synchronous. */
int rte_build_kdtree_gpu(rte_ctx* ctx, const float* pts, int64_t np, int64_t stride, int max_node,
                         int max_leaves, float* aabbs_out, int32_t* pids_out, int* n_leaf_out,
                         int where);
/* expandAabb (aabb.cu:27-38), host memory. */
void rte_expand_aabbs(float* aabbs, int n_leaf, float ratio);
/* hits[q] = smallest leaf whose box holds (PQ[q*stride+0..2]), 0 if none.
 * Replaces RT_Entry::Search -> OptiXRT::search (entry.cu:100-149,
 * raytracing.cpp:283-334, devicePrograms.cu:29-96). */
int rte_select_entry(rte_ctx* ctx, const float* PQ, int64_t nq, int stride,
                     int32_t* leaf_out, int where);

/* ---- search ------------------------------------------------------------ */
/* Greedy beam search of nq queries (row stride q_stride floats) from their
 * entry leaves.  ids_out/dists_out are nq x k.  Replaces
 * Graph::GraphSearch -> GraphSearchKernel (graph.cu:295-337,
 * graph_search.cuh:56-1964); distances are an addition.  stats_out (may be
 * NULL) receives nq x 3 int32: distance evaluations, graph rows read,
 * iterations. */
int rte_search(rte_ctx* ctx, const float* PQ, int64_t nq, int64_t q_stride,
               const int32_t* leaf, const rte_params* p, int32_t* ids_out,
               float* dists_out, int32_t* stats_out, int where);
/* Full-precision rerank (REORDER): for each query, the distances of its
 * n_cand candidate ids (e.g. the whole beam of rte_search with k = beam on
 * truncated PCA dimensions) are recomputed on the rerank base in the
 * original space, eight partial sums over dims l mod 8 added in order as the
 * reference's AVX loop does, and the k best by (dist, id) are written
 * (nq x k).  Ids outside the rerank base get +inf.  1 <= k <= n_cand <= 512,
 * d % 8 == 0, strides multiples of 4 floats.  Replaces the host
 * Graph::parallel_reorder (graph_helper.cu:146-213) called from
 * Graph::GraphSearch under REORDER (graph.cu:326-336). */
int rte_set_rerank_base(rte_ctx* ctx, const float* base, int64_t np, int d, int64_t stride,
                        int where);
int rte_rerank(rte_ctx* ctx, const float* Q, int64_t nq, int64_t q_stride, const int32_t* cand,
               int n_cand, int k, int32_t* ids_out, float* dists_out, int where);
/* PCA fit: FP64 mean and covariance / (n - 1) on the GPU (two passes,
 * fixed-order reductions), FP64 Householder + QL eigensolver on the host; R_out d x d
 * (host, column j = eigenvector of the j-th largest eigenvalue, largest-
 * magnitude component positive), mu_out d, eig_out d (may be NULL).  X is
 * n x d with row stride `stride` (host or device), n >= 2, d <= 4096.
 * Deterministic.  Replaces PCA::PCA (pca.cpp:4-58). */
int rte_pca_fit(rte_ctx* ctx, const float* X, int64_t n, int d, int64_t stride, float* R_out,
                float* mu_out, double* eig_out, int where);

/* out = (X - mu) R[:, :d_out] in FP64 (j ascending, multiply then add),
 * stored FP32: the base rotation of pca.cpp:122-139, bit-identical to that
 * host loop.  X n x d with row stride `stride` and out n x d_out are host or
 * device per `where`; R (d x d_out, row-major) and mu (d) are host arrays.
 * Synchronous. */
int rte_pca_rotate(rte_ctx* ctx, const float* X, int64_t n, int d, int64_t stride, const float* R,
                   const float* mu, int d_out, float* out, int where);

/* ---- graph construction ------------------------------------------------ */
/* Exact k nearest neighbours of every row of X (np x d, row stride `stride`)
 * among the rows of X, itself excluded, ordered by (squared L2 distance,
 * id): ids_out/dists_out np x k (dists may be NULL).  FP32 MFMA distances
 * |q|^2 + |x|^2 - 2 q.x; exact for integer-valued data with |x|^2 < 2^24.
 * d in {32, 64, 96, 128}, k <= 64, stride % 4 == 0.  No reference
 * counterpart: the reference loads prebuilt .nsw graphs and leaves graph
 * building as a TODO (graph.cu:114-139). */
int rte_knn(rte_ctx* ctx, const float* X, int64_t np, int d, int64_t stride, int k,
            int32_t* ids_out, float* dists_out, int where);

/* rte_knn for the rows [row0, row0 + nrows) of X only, still against all np
 * rows: ids_out/dists_out nrows x k.  A large build can run in pieces
 * (progress between calls, bounded output staging) or split its row ranges
 * across devices that each hold X.  No reference counterpart (as rte_knn). */
int rte_knn_rows(rte_ctx* ctx, const float* X, int64_t np, int d, int64_t stride, int64_t row0,
                 int64_t nrows, int k, int32_t* ids_out, float* dists_out, int where);

/* Search graph from kNN lists (np x kf, each row by (distance, id), e.g.
 * from rte_knn): candidate edges of u are u -> knn[u][j] (rank j) and, for
 * every v with knn[v][j] == u and j < rev_rank, u -> v (rank j); a pair keeps
 * its smallest (rank, forward before reverse); each row keeps the `degree`
 * candidates of smallest (rank, reverse, id), padded with np (graph_out
 * np x degree, ready for rte_set_graph).  kf, degree <= 64. */
int rte_graph_from_knn(rte_ctx* ctx, const int32_t* knn, int64_t np, int kf, int degree,
                       int rev_rank, int32_t* graph_out, int where);

/* Entry-space links (no reference counterpart; a graph-construction option):
 * the last r slots of every row of graph (np x degree, padded with np, host
 * or device, changed in place) become points drawn from the kd-tree leaves
 * l-window .. l+window around the row's own leaf l (leaf ids are in kd-tree
 * DFS order: a neighbourhood in the 3-D entry space), excluding the node
 * itself, its kept entries and each other (padding np after 8 draws).  On
 * data of many well-separated clusters the kNN graph has no edges between
 * clusters; these let a search whose entry seeds miss the query's cluster
 * reach the points that share the entry region with it.  Deterministic in
 * seed; needs rte_set_leaves (its pid table).  0 <= r <= min(degree, 32). */
int rte_graph_add_entry_links(rte_ctx* ctx, int32_t* graph, int64_t np, int degree, int r,
                              int window, uint64_t seed, int where);

/* Batched uneven search: per-query k and beam, one persistent launch with a
 * device work queue (longest beams first), captured as a hipGraph on first
 * use and replayed while the launch shape is unchanged; no host sync.  Per
 * query, on the device: k is clamped into [1, k_stride], beam becomes
 * pow2ceil(max(beam, k, 32)) capped at p->beam (0: 512), and k <= beam.
 * Outputs have row stride k_stride.  No reference counterpart; it
 * generalises GraphSearchKernel's launch (graph.cu:320-323) to mixed k and
 * beam.  p->k is ignored.
 * With k_per_q = beam_per_q = NULL the batch is uniform (every query uses
 * p->k and p->beam, k_stride >= p->k): the search kernel (after the leaf
 * sort with RTE_ORDER_LEAF) is launched directly, not through a graph (on
 * this stack a graph launch left the GPU idle for several microseconds on
 * either side of it), over p->grid workgroups (0: the resident maximum),
 * its waves taking queries from a work counter that the launch leaves at
 * zero; results equal rte_search's.  Batched calls on one context must not
 * overlap (they share the counter). */
int rte_search_batched(rte_ctx* ctx, const float* PQ, int64_t nq, int64_t q_stride,
                       const int32_t* leaf, const int32_t* k_per_q,
                       const int32_t* beam_per_q, const rte_params* p, int k_stride,
                       int32_t* ids_out, float* dists_out, int where);

/* rte_search_batched without any allocation (no reference counterpart; for
 * callers that cannot allocate device memory: code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * semantics and results as rte_search_batched, with every pointer a device
 * pointer; the kernels are launched directly on the context's stream (no
 * captured graph), and their work counters, per-query k/beam, query order
 * and leaf-sort buffers live in one static device array per GPU (3.8 MB,
 * held until the process ends), cleared by a kernel at the start of each
 * call.  Limits: nq <= RTE_NOALLOC_MAX_QUERIES, and with
 * RTE_ORDER_LEAF n_leaf (x 5 for an uneven batch) <= RTE_NOALLOC_MAX_KEYS;
 * beyond them the call fails with RTE_EINVAL.  Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (other streams, other
 * contexts, other threads) race and must be ordered by the caller.  The
 * rte_set_* calls that set the context up still allocate; make them first. */
#define RTE_NOALLOC_MAX_QUERIES 262144
#define RTE_NOALLOC_MAX_KEYS 81920
int rte_search_batched_noalloc(rte_ctx* ctx, const float* PQ, int64_t nq, int64_t q_stride,
                               const int32_t* leaf, const int32_t* k_per_q,
                               const int32_t* beam_per_q, const rte_params* p, int k_stride,
                               int32_t* ids_out, float* dists_out);

/* Autotuner for rte_search_batched: times the batch (device buffers) under
 * every launch shape -- wavefronts per query (wpq 1, 2, 4) x queries per
 * workgroup (wg_waves 1, 2, 4 at wpq 1) x beam storage (RTE_QUEUE_VGPR, and
 * RTE_QUEUE_LDS at wpq 1) x grid (RTE_GRID_PER_QUERY, RTE_GRID_DEFAULT,
 * RTE_GRID_RESIDENT2, RTE_GRID_BALANCED, each resolved from the occupancy
 * of the kernel launched) x query order (RTE_ORDER_GIVEN, RTE_ORDER_LEAF
 * with leaf entries) -- and every filter size in bloom_words[0..n_bloom)
 * (0 = the mode's default), and writes the fastest configuration to *best
 * (a copy of *p with wpq, wg_waves, queue, grid, order and bloom_words set)
 * and its time to *best_ms.  Each shape is timed over three launches (the
 * mean); the shapes within 5% of the fastest mean (at most six) are then
 * timed again in nine interleaved rounds, one launch of each per round, and
 * the lowest median wins (*best_ms is that median).  A shape that fails is
 * skipped; the call fails only when every shape did (with the first
 * failure).  Launch shape and order never change results; the filter size
 * can, so callers pass only sizes they accept.  ids_out/dists_out receive
 * the chosen configuration's results. */
int rte_autotune_batched(rte_ctx* ctx, const float* PQ, int64_t nq, int64_t q_stride,
                         const int32_t* leaf, const int32_t* k_per_q,
                         const int32_t* beam_per_q, const rte_params* p, int k_stride,
                         int32_t* ids_out, float* dists_out, const int* bloom_words,
                         int n_bloom, rte_params* best, float* best_ms);

/* Whole query path on device buffers (graph.cu:254-293 Graph::Search):
 * rte_project into PQ_scratch (nq x d_out), rte_select_entry into
 * leaf_scratch when p->algo == 1 (the random entries of the other algo
 * values use no leaves), then rte_search; same results as those calls. */
int rte_query(rte_ctx* ctx, const float* Q, int64_t nq, const rte_params* p,
              float* PQ_scratch, int32_t* leaf_scratch, int32_t* ids_out,
              float* dists_out);

/* ---- learned entry predictor (no reference counterpart) ---------------- */
/* MLP d_in -> h1 -> h2 -> n_leaf, ReLU, bf16 weights, FP32 accumulation. */
int rte_mlp_create(rte_ctx* ctx, int d_in, int h1, int h2, int n_out, uint64_t seed);
/* One Adam step (beta1 = momentum, beta2 = 0.999) on a device batch (X: n x
 * d_in f32, y: n int32 labels); writes the mean cross-entropy loss to *loss
 * (host) when loss is not NULL. */
int rte_mlp_train_step(rte_ctx* ctx, const float* X, const int32_t* y, int n, float lr,
                       float momentum, float* loss);
/* Predicted leaf and softmax confidence per query (device buffers).
 * leaf_out[q] falls back to fallback[q] (the kd-tree leaf) when the
 * softmax maximum is below min_conf. */
int rte_mlp_predict(rte_ctx* ctx, const float* X, int64_t n, const int32_t* fallback,
                    float min_conf, int32_t* leaf_out, float* conf_out);
/* Entry leaves of nq PCA queries (device buffers, row stride `stride` >=
 * max(3, d_in)): the network's leaf, or the kd-tree leaf of rte_select_entry
 * when the softmax maximum is below min_conf.  Same results as
 * rte_select_entry followed by rte_mlp_predict; the kd-tree traversal runs
 * in workgroups of the hidden layers' launch, beside the network, on the
 * context's stream.  Needs rte_set_leaves and rte_mlp_create. */
int rte_predict_entry(rte_ctx* ctx, const float* PQ, int64_t nq, int stride, float min_conf,
                      int32_t* leaf_out, float* conf_out);
/* Input normalisation x' = (x - mean) * inv_std per feature (host arrays of
 * d_in), applied in FP32 before the bf16 rounding; identity until set. */
int rte_mlp_set_input_norm(rte_ctx* ctx, const float* mean, const float* inv_std);
/* Weights in/out as FP32 host arrays (W1 d_in x h1, b1, W2 h1 x h2, b2,
 * W3 h2 x n_out, b3), for checkpoints and the gradient check. */
int rte_mlp_get(rte_ctx* ctx, float* W1, float* b1, float* W2, float* b2, float* W3,
                float* b3);
int rte_mlp_set(rte_ctx* ctx, const float* W1, const float* b1, const float* W2,
                const float* b2, const float* W3, const float* b3);
/* Gradients of the mean cross-entropy of one device batch with respect to
 * the FP32 master weights (host outputs, same shapes as rte_mlp_get). */
int rte_mlp_grad(rte_ctx* ctx, const float* X, const int32_t* y, int n, float* gW1,
                 float* gb1, float* gW2, float* gb2, float* gW3, float* gb3,
                 float* loss);

/* ---- misc -------------------------------------------------------------- */
/* auto_tune_bloom.h:67-89 hash_parameter, used to size the TUNED filter. */
void rte_hash_parameter(int nc, int* hash_len, int* bit, int* hash);
/* Build configuration: 1 if (beam, batch, dims-per-lane) has a compiled
 * search kernel. */
int rte_search_supported(int beam, int batch, int d);
const char* rte_version(void);

#ifdef __cplusplus
}
#endif
#endif
