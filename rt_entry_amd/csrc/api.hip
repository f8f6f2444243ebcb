// C ABI of the MI355X query path (include/rt_entry.h): context, device
// buffers, argument checks and launches.  Host memory arguments are staged
// through context-owned device buffers; device arguments are used in place.
#include "rt_entry.h"

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "ctx.h"
#include "entry.h"
#include "kdtree.h"
#include "project.h"
#include "rerank.h"
#include "knn.h"
#include "links.h"
#include "search.h"

namespace {

void hash_parameter(int nc, int* hash_len, int* bit, int* hash);

}  // namespace

int rte_fail(rte_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

namespace {

#define fail rte_fail

int hip_fail(rte_ctx* c, hipError_t e, const char* what) {
  if (e == hipErrorOutOfMemory)
    return fail(c, RTE_ENOMEM, "%s: %s", what, hipGetErrorString(e));
  return fail(c, RTE_EHIP, "%s: %s", what, hipGetErrorString(e));
}

#define HIPCHK(ctx, expr, what)                      \
  do {                                               \
    hipError_t e_ = (expr);                          \
    if (e_ != hipSuccess) return hip_fail(ctx, e_, what); \
  } while (0)

bool is_pow2(int x) { return x > 0 && (x & (x - 1)) == 0; }
int pow2_ceil(int x) {
  int p = 1;
  while (p < x) p <<= 1;
  return p;
}
int ilog2(int x) {
  int s = 0;
  while ((1 << s) < x) ++s;
  return s;
}

void hash_parameter(int nc, int* hash_len, int* bit, int* hash) {
  // auto_tune_bloom.h:67-89: n = 12.897 nc; the smallest m = 32 * 2^i
  // (i = 1..11) whose false-positive rate with k = round(ln2 m/n) is < 3%.
  const double n = (double)nc * 12.897;
  const int maxn = 11;
  for (int i = 1; i <= maxn; ++i) {
    const double m = 32.0 * (double)(1 << i);
    const double r = m / n;
    const double k = std::round(std::log(2.0) * r);
    const double p = std::pow(1.0 - std::exp(-k / r), k);
    if (p < 0.03) {
      *hash_len = 1 << i;
      *bit = i;
      *hash = (int)std::round(k);
      return;
    }
    if (i == maxn) *hash_len = 1 << i;  // the reference's unbraced if
    *bit = i;
    *hash = (int)std::round(k);
  }
}

// Resolve the filter configuration for one beam.
void bloom_for(const rte_params* p, int beam, int* words, int* hashes) {
  // hash count per mode is fixed in the kernels: 3 (graph_search.cuh:5) or
  // auto_tune_bloom.h's k (7 for every beam 32..512)
  *hashes = p->mode == RTE_MODE_TUNED ? 7 : 3;
  if (p->bloom_words > 0) {
    *words = p->bloom_words;
    return;
  }
  if (p->mode == RTE_MODE_TUNED) {
    int hl, bit, hh;
    hash_parameter(beam, &hl, &bit, &hh);
    *words = hl;
    return;
  }
  *words = 512;  // graph_search.cuh:4
}

int check_params(rte_ctx* c, const rte_params* p, bool per_query) {
  if (!p) return fail(c, RTE_EINVAL, "params is NULL");
  if (!per_query) {
    if (p->beam < 1 || p->beam > 512)
      return fail(c, RTE_EINVAL, "beam %d outside 1..512", p->beam);
    if (p->k < 1 || p->k > pow2_ceil(p->beam))
      return fail(c, RTE_EINVAL, "k %d must be in 1..beam", p->k);
  }
  if (p->width < 1 || p->width > rte::kMaxWidth)
    return fail(c, RTE_EINVAL, "width %d outside 1..%d", p->width, rte::kMaxWidth);
  if (p->max_iter < 0) return fail(c, RTE_EINVAL, "max_iter < 0");
  if (p->mode != RTE_MODE_REF && p->mode != RTE_MODE_TUNED)
    return fail(c, RTE_EINVAL, "unknown mode %d", p->mode);
  if (p->algo < 0 || p->algo > 2) return fail(c, RTE_EINVAL, "unknown algo %d", p->algo);
  if (p->bloom_words != 0 &&
      (!is_pow2(p->bloom_words) || p->bloom_words < 2 || p->bloom_words > 2048))
    return fail(c, RTE_EINVAL, "bloom_words must be a power of two in 2..2048");
  if (p->bloom_hashes != 0 && p->bloom_hashes != (p->mode == RTE_MODE_TUNED ? 7 : 3))
    return fail(c, RTE_EUNSUP, "bloom_hashes must be 0 or %d in this mode",
                p->mode == RTE_MODE_TUNED ? 7 : 3);
  return RTE_OK;
}

// device pointer for an input that may live on the host
template <typename T>
int stage_in(rte_ctx* c, DevBuf& buf, const T* src, size_t count, int where, const T** dev) {
  if (where == RTE_DEVICE) {
    *dev = src;
    return RTE_OK;
  }
  HIPCHK(c, buf.reserve(count * sizeof(T)), "hipMalloc(staging)");
  HIPCHK(c, hipMemcpyAsync(buf.p, src, count * sizeof(T), hipMemcpyHostToDevice, c->stream),
         "hipMemcpyAsync H2D");
  *dev = buf.as<T>();
  return RTE_OK;
}

template <typename T>
int stage_out(rte_ctx* c, DevBuf& buf, T* dst, size_t count, int where, T** dev) {
  if (where == RTE_DEVICE || dst == nullptr) {
    *dev = dst;
    return RTE_OK;
  }
  HIPCHK(c, buf.reserve(count * sizeof(T)), "hipMalloc(staging)");
  *dev = buf.as<T>();
  return RTE_OK;
}

template <typename T>
int finish_out(rte_ctx* c, T* host, const T* dev, size_t count, int where) {
  if (where == RTE_DEVICE || host == nullptr) return RTE_OK;
  HIPCHK(c, hipMemcpyAsync(host, dev, count * sizeof(T), hipMemcpyDeviceToHost, c->stream),
         "hipMemcpyAsync D2H");
  HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  return RTE_OK;
}

int set_device(rte_ctx* c) {
  HIPCHK(c, hipSetDevice(c->device), "hipSetDevice");
  return RTE_OK;
}

}  // namespace

static hipError_t set_sort_kernel_attributes();

extern "C" {

int rte_create(rte_ctx** out, int device) {
  if (!out) return RTE_EINVAL;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return RTE_EHIP;
  if (device < 0 || device >= n) return RTE_EINVAL;
  rte_ctx* c = new rte_ctx();
  c->device = device;
  // a blocking stream: ordered with the null stream, so work the caller
  // queues there (torch's default stream) and the context's kernels see each
  // other's results without explicit events
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->own) != hipSuccess) {
    delete c;
    return RTE_EHIP;
  }
  if (set_sort_kernel_attributes() != hipSuccess) {
    (void)hipStreamDestroy(c->own);
    delete c;
    return RTE_EHIP;
  }
  c->stream = c->own;
  *out = c;
  return RTE_OK;
}

int rte_destroy(rte_ctx* c) {
  if (!c) return RTE_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->own) (void)hipStreamDestroy(c->own);
  delete c;
  return RTE_OK;
}

const char* rte_last_error(const rte_ctx* c) { return c ? c->err.c_str() : "null context"; }

int rte_set_stream(rte_ctx* c, void* stream) {
  if (!c) return RTE_EINVAL;
  c->stream = stream ? (hipStream_t)stream : c->own;
  return RTE_OK;
}

void* rte_get_stream(rte_ctx* c) { return c ? (void*)c->stream : nullptr; }

int rte_synchronize(rte_ctx* c) {
  if (!c) return RTE_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
  return RTE_OK;
}

int rte_set_base(rte_ctx* c, const float* base, int64_t np, int d, int64_t stride, int where) {
  if (!c) return RTE_EINVAL;
  // the search forms row byte offsets as 32-bit id x 32-bit row bytes
  if (!base || np <= 0 || np >= 0x7fffffff || d <= 0 || stride < d || stride >= (1 << 30))
    return fail(c, RTE_EINVAL, "bad base shape np=%lld d=%d stride=%lld", (long long)np, d,
                (long long)stride);
  if (d > 960) return fail(c, RTE_EUNSUP, "d=%d above 960", d);
  if (int r = set_device(c)) return r;
  if (where == RTE_HOST) {
    const size_t bytes = (size_t)np * stride * sizeof(float);
    HIPCHK(c, c->base_own.reserve(bytes), "hipMalloc(base)");
    HIPCHK(c, hipMemcpyAsync(c->base_own.p, base, bytes, hipMemcpyHostToDevice, c->stream),
           "copy base");
    HIPCHK(c, hipStreamSynchronize(c->stream), "sync base");
    c->base = c->base_own.as<float>();
  } else {
    c->base_own.release();
    c->base = base;
  }
  c->np = np;
  c->d = d;
  c->stride = stride;
  return RTE_OK;
}

int rte_set_rerank_base(rte_ctx* c, const float* base, int64_t np, int d, int64_t stride,
                        int where) {
  if (!c) return RTE_EINVAL;
  if (!base || np <= 0 || np >= 0x7fffffff || d <= 0 || d % 8 != 0 || stride < d ||
      stride % 4 != 0)
    return fail(c, RTE_EINVAL, "bad rerank base np=%lld d=%d stride=%lld (d %% 8, stride %% 4)",
                (long long)np, d, (long long)stride);
  if (int r = set_device(c)) return r;
  if (where == RTE_HOST) {
    const size_t bytes = (size_t)np * stride * sizeof(float);
    HIPCHK(c, c->rr_own.reserve(bytes), "hipMalloc(rerank base)");
    HIPCHK(c, hipMemcpyAsync(c->rr_own.p, base, bytes, hipMemcpyHostToDevice, c->stream),
           "copy rerank base");
    HIPCHK(c, hipStreamSynchronize(c->stream), "sync rerank base");
    c->rr_base = c->rr_own.as<float>();
  } else {
    if (((uintptr_t)base & 15) != 0) return fail(c, RTE_EINVAL, "rerank base not 16-byte aligned");
    c->rr_own.release();
    c->rr_base = base;
  }
  c->rr_np = np;
  c->rr_d = d;
  c->rr_stride = stride;
  return RTE_OK;
}

int rte_rerank(rte_ctx* c, const float* Q, int64_t nq, int64_t q_stride, const int32_t* cand,
               int n_cand, int k, int32_t* ids_out, float* dists_out, int where) {
  if (!c) return RTE_EINVAL;
  if (!c->rr_base) return fail(c, RTE_ESTATE, "rte_set_rerank_base not called");
  if ((nq > 0 && (!Q || !cand || !ids_out)) || nq < 0 || q_stride < c->rr_d || q_stride % 4 != 0)
    return fail(c, RTE_EINVAL, "bad rerank args (q_stride %lld, d %d)", (long long)q_stride,
                c->rr_d);
  if (n_cand < 1 || n_cand > rte::kRerankMaxCand || k < 1 || k > n_cand)
    return fail(c, RTE_EINVAL, "rerank needs 1 <= k <= n_cand <= %d (k %d, n_cand %d)",
                rte::kRerankMaxCand, k, n_cand);
  if (nq == 0) return RTE_OK;
  if (int r = set_device(c)) return r;
  const float* dq;
  const int32_t* dc;
  int32_t* di;
  float* dd = nullptr;
  if (int r = stage_in(c, c->s_in, Q, (size_t)nq * q_stride, where, &dq)) return r;
  if (where == RTE_DEVICE && ((uintptr_t)Q & 15) != 0)
    return fail(c, RTE_EINVAL, "queries not 16-byte aligned");
  if (int r = stage_in(c, c->s_aux, cand, (size_t)nq * n_cand, where, &dc)) return r;
  if (int r = stage_out(c, c->s_out, ids_out, (size_t)nq * k, where, &di)) return r;
  if (dists_out)
    if (int r = stage_out(c, c->s_aux2, dists_out, (size_t)nq * k, where, &dd)) return r;
  HIPCHK(c, rte::launch_rerank(c->rr_base, c->rr_np, c->rr_stride, c->rr_d, dq, nq, q_stride, dc,
                               n_cand, k, di, dd, c->stream),
         "rerank launch");
  if (int r = finish_out(c, ids_out, di, (size_t)nq * k, where)) return r;
  if (dists_out)
    if (int r = finish_out(c, dists_out, dd, (size_t)nq * k, where)) return r;
  return RTE_OK;
}

int rte_knn_rows(rte_ctx* c, const float* X, int64_t np, int d, int64_t stride, int64_t row0,
                 int64_t nrows, int k, int32_t* ids_out, float* dists_out, int where) {
  if (!c) return RTE_EINVAL;
  if (!X || !ids_out || np < 2 || np >= 0x7fffffff || stride < d || stride % 4 != 0)
    return fail(c, RTE_EINVAL, "bad knn arguments (np %lld, d %d, stride %lld)", (long long)np, d,
                (long long)stride);
  if (row0 < 0 || nrows < 1 || row0 + nrows > np)
    return fail(c, RTE_EINVAL, "knn rows [%lld, %lld) outside [0, %lld)", (long long)row0,
                (long long)(row0 + nrows), (long long)np);
  if (d != 32 && d != 64 && d != 96 && d != 128)
    return fail(c, RTE_EUNSUP, "knn kernel compiled for d in {32, 64, 96, 128}, not %d", d);
  if (k < 1 || k > rte::kKnnK || k >= np)
    return fail(c, RTE_EINVAL, "knn needs 1 <= k <= %d and k < np", rte::kKnnK);
  if (int r = set_device(c)) return r;
  const float* dx;
  if (int r = stage_in(c, c->s_in, X, (size_t)np * stride, where, &dx)) return r;
  if (where == RTE_DEVICE && ((uintptr_t)X & 15) != 0)
    return fail(c, RTE_EINVAL, "X not 16-byte aligned");
  int32_t* di;
  float* dd = nullptr;
  if (int r = stage_out(c, c->s_out, ids_out, (size_t)nrows * k, where, &di)) return r;
  if (dists_out)
    if (int r = stage_out(c, c->s_aux2, dists_out, (size_t)nrows * k, where, &dd)) return r;
  DevBuf norms;  // freed on return (hipFree waits for the device)
  HIPCHK(c, norms.reserve((size_t)np * sizeof(float)), "hipMalloc(norms)");
  HIPCHK(c, rte::launch_row_norms(dx, np, d, stride, norms.as<float>(), c->stream), "norms");
  // rows in launches of 128k (short kernels)
  const int64_t step = 1 << 17;
  for (int64_t r0 = row0; r0 < row0 + nrows; r0 += step) {
    const int64_t n = std::min<int64_t>(step, row0 + nrows - r0);
    HIPCHK(c, rte::launch_knn(dx, np, d, stride, norms.as<float>(), r0, n, k, di + (r0 - row0) * k,
                              dd ? dd + (r0 - row0) * k : nullptr, c->stream),
           "knn launch");
  }
  if (int r = finish_out(c, ids_out, di, (size_t)nrows * k, where)) return r;
  if (dists_out)
    if (int r = finish_out(c, dists_out, dd, (size_t)nrows * k, where)) return r;
  HIPCHK(c, hipStreamSynchronize(c->stream), "sync knn");  // norms is freed on return
  return RTE_OK;
}

int rte_knn(rte_ctx* c, const float* X, int64_t np, int d, int64_t stride, int k,
            int32_t* ids_out, float* dists_out, int where) {
  return rte_knn_rows(c, X, np, d, stride, 0, np, k, ids_out, dists_out, where);
}

int rte_graph_from_knn(rte_ctx* c, const int32_t* knn, int64_t np, int kf, int degree,
                       int rev_rank, int32_t* graph_out, int where) {
  if (!c) return RTE_EINVAL;
  if (!knn || !graph_out || np < 1 || np >= 0x7fffffff || kf < 1 || kf > 64 || degree < 1 ||
      degree > 64 || rev_rank < 0)
    return fail(c, RTE_EINVAL, "bad graph arguments (kf %d, degree %d <= 64)", kf, degree);
  if (int r = set_device(c)) return r;
  const int32_t* dk;
  int32_t* dg;
  if (int r = stage_in(c, c->s_aux, knn, (size_t)np * kf, where, &dk)) return r;
  if (int r = stage_out(c, c->s_out, graph_out, (size_t)np * degree, where, &dg)) return r;
  DevBuf work;  // freed on return (hipFree waits for the device)
  HIPCHK(c, work.reserve(rte::graph_from_knn_work_bytes(np, kf, rev_rank)), "hipMalloc(graph work)");
  HIPCHK(c, rte::launch_graph_from_knn(dk, np, kf, degree, rev_rank, dg, work.p, c->stream),
         "graph launch");
  if (int r = finish_out(c, graph_out, dg, (size_t)np * degree, where)) return r;
  HIPCHK(c, hipStreamSynchronize(c->stream), "sync graph");
  return RTE_OK;
}

int rte_graph_add_entry_links(rte_ctx* c, int32_t* graph, int64_t np, int degree, int r,
                              int window, uint64_t seed, int where) {
  if (!c) return RTE_EINVAL;
  if (!c->pids || c->n_leaf <= 0) return fail(c, RTE_ESTATE, "rte_set_leaves not called");
  if (!graph || np <= 0 || np >= 0x7fffffff || degree < 1 || degree > 1024)
    return fail(c, RTE_EINVAL, "bad graph shape np=%lld degree=%d", (long long)np, degree);
  if (r < 0 || r > degree || r > rte::kMaxLinks || window < 0)
    return fail(c, RTE_EINVAL, "links need 0 <= r <= min(degree, %d) and window >= 0 (r %d)",
                rte::kMaxLinks, r);
  if (r == 0) return RTE_OK;
  if (int e = set_device(c)) return e;
  DevBuf gbuf, leaf_of, size;
  int32_t* dg = graph;
  const size_t bytes = (size_t)np * degree * sizeof(int32_t);
  if (where == RTE_HOST) {
    HIPCHK(c, gbuf.reserve(bytes), "hipMalloc(graph)");
    HIPCHK(c, hipMemcpyAsync(gbuf.p, graph, bytes, hipMemcpyHostToDevice, c->stream), "copy graph");
    dg = gbuf.as<int32_t>();
  }
  HIPCHK(c, leaf_of.reserve((size_t)np * sizeof(int32_t)), "hipMalloc(leaf index)");
  HIPCHK(c, size.reserve((size_t)c->n_leaf * sizeof(int32_t)), "hipMalloc(leaf sizes)");
  HIPCHK(c, rte::launch_entry_links(dg, np, degree, r, window, seed, c->pids, c->n_leaf, c->cap,
                                    leaf_of.as<int32_t>(), size.as<int32_t>(), c->stream),
         "entry links");
  if (where == RTE_HOST)
    HIPCHK(c, hipMemcpyAsync(graph, dg, bytes, hipMemcpyDeviceToHost, c->stream), "copy graph");
  HIPCHK(c, hipStreamSynchronize(c->stream), "sync links");
  return RTE_OK;
}

__global__ void repack_graph_kernel(const int32_t* __restrict__ src, int64_t np, int degree,
                                    int stride, int32_t* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np * stride) return;
  const int64_t row = i / stride;
  const int col = (int)(i - row * stride);
  dst[i] = col < degree ? src[row * degree + col] : (int32_t)np;
}

int rte_set_graph(rte_ctx* c, const int32_t* rows, int64_t np, int degree, int where) {
  if (!c) return RTE_EINVAL;
  if (!rows || np <= 0 || np >= 0x7fffffff || degree < 1 || degree > 64)
    return fail(c, RTE_EINVAL, "bad graph shape np=%lld degree=%d (1..64)", (long long)np, degree);
  if (int r = set_device(c)) return r;
  const int D = pow2_ceil(degree);
  const int shift = ilog2(D);
  if (where == RTE_HOST) {
    for (int64_t i = 0; i < np * degree; ++i)
      if (rows[i] < 0 || rows[i] > np)
        return fail(c, RTE_EINVAL, "graph id %d at %lld outside [0,np]", rows[i], (long long)i);
    std::vector<int32_t> padded;
    const int32_t* src = rows;
    if (D != degree) {
      padded.assign((size_t)np * D, (int32_t)np);
      for (int64_t r = 0; r < np; ++r)
        memcpy(&padded[(size_t)r * D], rows + r * degree, sizeof(int32_t) * degree);
      src = padded.data();
    }
    const size_t bytes = (size_t)np * D * sizeof(int32_t);
    HIPCHK(c, c->graph_own.reserve(bytes), "hipMalloc(graph)");
    HIPCHK(c, hipMemcpyAsync(c->graph_own.p, src, bytes, hipMemcpyHostToDevice, c->stream),
           "copy graph");
    HIPCHK(c, hipStreamSynchronize(c->stream), "sync graph");
    c->graph = c->graph_own.as<int32_t>();
  } else if (D != degree) {
    const size_t bytes = (size_t)np * D * sizeof(int32_t);
    HIPCHK(c, c->graph_own.reserve(bytes), "hipMalloc(graph)");
    const int64_t tot = np * D;
    hipLaunchKernelGGL(repack_graph_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                       c->stream, rows, np, degree, D, c->graph_own.as<int32_t>());
    HIPCHK(c, hipGetLastError(), "repack graph");
    c->graph = c->graph_own.as<int32_t>();
  } else {
    c->graph_own.release();
    c->graph = rows;
  }
  c->graph_np = np;
  c->degree = degree;
  c->shift = shift;
  return RTE_OK;
}

int rte_set_pca(rte_ctx* c, const float* R, const float* mu, int d, int r_cols, int d_out,
                int where) {
  if (!c) return RTE_EINVAL;
  if (!R || !mu || d <= 0 || d > 4096 || r_cols < d_out || d_out <= 0)
    return fail(c, RTE_EINVAL, "bad pca shape d=%d r_cols=%d d_out=%d", d, r_cols, d_out);
  if (int r = set_device(c)) return r;
  HIPCHK(c, c->rot.reserve((size_t)d * d_out * sizeof(float)), "hipMalloc(rot)");
  HIPCHK(c, c->bias.reserve((size_t)d_out * sizeof(float)), "hipMalloc(bias)");
  const hipMemcpyKind kind = where == RTE_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
  // keep the first d_out columns (graph.cu:196 CopyHostToDevice(..., DIM))
  HIPCHK(c, hipMemcpy2DAsync(c->rot.p, d_out * sizeof(float), R, r_cols * sizeof(float),
                             d_out * sizeof(float), d, kind, c->stream),
         "copy rotation");
  DevBuf mu_dev;
  HIPCHK(c, mu_dev.reserve(d * sizeof(float)), "hipMalloc(mu)");
  HIPCHK(c, hipMemcpyAsync(mu_dev.p, mu, d * sizeof(float), kind, c->stream), "copy mu");
  HIPCHK(c, rte::launch_mu_r(mu_dev.as<float>(), c->rot.as<float>(), d, d_out, c->bias.as<float>(),
                             c->stream),
         "mu.R");
  HIPCHK(c, hipStreamSynchronize(c->stream), "sync pca");
  c->pca_d = d;
  c->pca_out = d_out;
  return RTE_OK;
}

int rte_project(rte_ctx* c, const float* Q, int64_t nq, float* PQ, int where) {
  if (!c) return RTE_EINVAL;
  if (!c->pca_d) return fail(c, RTE_ESTATE, "rte_set_pca not called");
  if ((nq > 0 && (!Q || !PQ)) || nq < 0) return fail(c, RTE_EINVAL, "bad projection arguments");
  if (nq == 0) return RTE_OK;
  if (int r = set_device(c)) return r;
  const float* dq;
  float* dpq;
  if (int r = stage_in(c, c->s_in, Q, (size_t)nq * c->pca_d, where, &dq)) return r;
  if (int r = stage_out(c, c->s_out, PQ, (size_t)nq * c->pca_out, where, &dpq)) return r;
  HIPCHK(c, rte::launch_project(dq, nq, c->pca_d, c->rot.as<float>(), c->pca_out,
                                c->bias.as<float>(), dpq, c->stream),
         "project");
  return finish_out(c, PQ, dpq, (size_t)nq * c->pca_out, where);
}

int rte_build_kdtree(const float* pts3, int64_t np, int max_node, int max_leaves,
                     float* aabbs_out, int32_t* pids_out, int* n_leaf_out) {
  if (!pts3 || !aabbs_out || !pids_out || !n_leaf_out || np <= 0 || max_node <= 0 ||
      max_leaves <= 0)
    return RTE_EINVAL;
  const int r = rte::build_kdtree(pts3, np, max_node, max_leaves, aabbs_out, pids_out, n_leaf_out);
  if (r == -2) return RTE_EINVAL;
  if (r == -3) return RTE_EBUILD;
  return r == 0 ? RTE_OK : RTE_EINVAL;
}

void rte_expand_aabbs(float* aabbs, int n_leaf, float ratio) {
  if (aabbs && n_leaf > 0) rte::expand_aabbs(aabbs, n_leaf, ratio);
}

int rte_set_leaves(rte_ctx* c, const float* aabbs, const int32_t* pids, int n_leaf, int cap,
                   int where) {
  if (!c) return RTE_EINVAL;
  if (!aabbs || !pids || n_leaf <= 0 || cap <= 0)
    return fail(c, RTE_EINVAL, "bad leaves n_leaf=%d cap=%d", n_leaf, cap);
  if (int r = set_device(c)) return r;
  std::vector<float> hb;
  const float* boxes = aabbs;
  if (where == RTE_DEVICE) {
    hb.resize((size_t)n_leaf * 6);
    HIPCHK(c, hipMemcpy(hb.data(), aabbs, hb.size() * sizeof(float), hipMemcpyDeviceToHost),
           "read boxes");
    boxes = hb.data();
  }
  std::vector<rte::BoxNode> nodes;
  rte::build_hierarchy(boxes, n_leaf, nodes);
  HIPCHK(c, c->nodes.reserve(nodes.size() * sizeof(rte::BoxNode)), "hipMalloc(nodes)");
  HIPCHK(c, hipMemcpyAsync(c->nodes.p, nodes.data(), nodes.size() * sizeof(rte::BoxNode),
                           hipMemcpyHostToDevice, c->stream),
         "copy nodes");
  if (where == RTE_HOST) {
    const size_t bytes = (size_t)n_leaf * cap * sizeof(int32_t);
    HIPCHK(c, c->pids_own.reserve(bytes), "hipMalloc(pids)");
    HIPCHK(c, hipMemcpyAsync(c->pids_own.p, pids, bytes, hipMemcpyHostToDevice, c->stream),
           "copy pids");
    c->pids = c->pids_own.as<int32_t>();
  } else {
    c->pids_own.release();
    c->pids = pids;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream), "sync leaves");
  c->n_nodes = (int)nodes.size();
  c->n_leaf = n_leaf;
  c->cap = cap;
  return RTE_OK;
}

int rte_select_entry(rte_ctx* c, const float* PQ, int64_t nq, int stride, int32_t* leaf_out,
                     int where) {
  if (!c) return RTE_EINVAL;
  if (!c->n_nodes) return fail(c, RTE_ESTATE, "rte_set_leaves not called");
  if ((nq > 0 && (!PQ || !leaf_out)) || nq < 0 || stride < 3)
    return fail(c, RTE_EINVAL, "bad entry args");
  if (nq == 0) return RTE_OK;
  if (int r = set_device(c)) return r;
  const float* dpq;
  int32_t* dl;
  if (int r = stage_in(c, c->s_in, PQ, (size_t)nq * stride, where, &dpq)) return r;
  if (int r = stage_out(c, c->s_out, leaf_out, (size_t)nq, where, &dl)) return r;
  HIPCHK(c, rte::launch_select_entry(c->nodes.as<rte::BoxNode>(), c->n_nodes, dpq, nq, stride, dl,
                                     c->stream),
         "select_entry");
  return finish_out(c, leaf_out, dl, (size_t)nq, where);
}

struct BatchPrep {
  const int32_t* k_in;
  const int32_t* beam_in;
  int64_t nq;
  int k_stride;
  int beam_cap;
};

// counter[0] = work counter, counter[1..5] = per-beam-class counts, then
// counter[8..12] = scatter cursors; one thread, vector stores (the address
// depends on threadIdx)
__global__ void batch_reset_kernel(int64_t* counter) {
  for (int i = 0; i < 16; ++i) counter[threadIdx.x + i] = 0;
}

// k into [1, min(k_stride, beam)], beam = pow2ceil(max(beam, k, 32)) capped;
// counts per beam class (32,64,...,512 -> 0..4), one atomic per wave and
// class (same-address device atomics serialise)
__global__ void batch_prep_kernel(BatchPrep b, int32_t* __restrict__ k_out,
                                  int32_t* __restrict__ beam_out, int64_t* counter) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool in = q < b.nq;
  int cl = -1;
  if (in) {
    int k = b.k_in[q];
    k = k < 1 ? 1 : (k > b.k_stride ? b.k_stride : k);
    int bm = b.beam_in[q];
    bm = bm > k ? bm : k;
    bm = bm > 32 ? bm : 32;
    int p2 = 32;
    while (p2 < bm && p2 < b.beam_cap) p2 <<= 1;
    if (k > p2) k = p2;
    k_out[q] = k;
    beam_out[q] = p2;
    cl = (31 - __clz(p2)) - 5;
  }
  if (!counter) return;  // the keyed order counts for itself
  const int lane = threadIdx.x & 63;
  for (int c = 0; c < 5; ++c) {
    const uint64_t m = __ballot(cl == c);
    if (m && lane == __ffsll((unsigned long long)m) - 1)
      atomicAdd(reinterpret_cast<unsigned long long*>(counter + 1 + c),
                (unsigned long long)__popcll(m));
  }
}

// RTE_ORDER_LEAF: a counting sort of the batch by key = entry leaf, after
// the beam class (longest first) in an uneven batch.  Leaves are numbered in
// kd-tree DFS order, so neighbouring keys are neighbouring regions of the
// PCA space, whose searches gather many of the same rows.
struct KeySort {
  const int32_t* leaf;    // entry leaves
  const int32_t* beam;    // effective beams (uneven batch) or null
  int64_t nq;
  int n_leaf;
  int n_keys;             // n_leaf, or 5 * n_leaf with beam classes
  int32_t* hist;          // n_keys counts, zero on entry, left zero by the scan
  int32_t* cursor;        // n_keys scatter cursors
  int32_t* order;         // nq query ids by key
  int64_t* seg;           // 16 * segments work counters, reset by the scan
  int segments;
};

// key of a query with entry leaf l and effective beam b (b unused without
// beam classes)
__device__ __forceinline__ int key_from(const KeySort& k, int l, int b) {
  if (l < 0 || l >= k.n_leaf) l = 0;  // the search's own clamp of a bad leaf
  if (!k.beam) return l;
  const int cl = (31 - __clz(b)) - 5;  // beams 32..512 -> 0..4
  return (4 - cl) * k.n_leaf + l;
}

__device__ __forceinline__ int sort_key(const KeySort& k, int64_t q) {
  return key_from(k, k.leaf ? k.leaf[q] : 0, k.beam ? k.beam[q] : 32);
}

__global__ void key_hist_kernel(KeySort k) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < k.nq) atomicAdd(&k.hist[sort_key(k, q)], 1);
}

// One workgroup of 1024 threads: exclusive prefix sums of the histogram into
// the cursors, in passes of kScanChunk keys staged in LDS.  A pass loads its
// keys with coalesced reads (all issued before any store, which would make
// each later load wait), thread t sums the consecutive keys
// [t*per, (t+1)*per) of the pass, a workgroup scan combines the sums, and
// the prefixes leave with coalesced writes that also zero the histogram.
constexpr int kScanChunk = 8192;

// The counts vals[0, nc) of one pass (in LDS) become exclusive prefixes
// plus carry, in place; returns carry plus the pass's total.  All 1024
// threads of the workgroup call it (it holds barriers); thread t sums the
// consecutive keys [t*per, (t+1)*per), a Hillis-Steele scan combines the
// sums.
__device__ __forceinline__ int32_t scan_pass(int32_t* vals, int32_t (*part)[1024], int nc,
                                             int32_t carry, int t) {
  const int per = (nc + 1023) / 1024;
  const int k0 = min(nc, t * per), k1 = min(nc, k0 + per);
  int32_t sum = 0;
  for (int i = k0; i < k1; ++i) sum += vals[i];
  part[0][t] = sum;
  __syncthreads();
  int src = 0;
  for (int off = 1; off < 1024; off <<= 1) {  // inclusive Hillis-Steele scan
    const int32_t v = part[src][t] + (t >= off ? part[src][t - off] : 0);
    part[src ^ 1][t] = v;
    __syncthreads();
    src ^= 1;
  }
  int32_t run = carry + part[src][t] - sum;  // exclusive prefix of this thread's keys
  for (int i = k0; i < k1; ++i) {
    const int32_t h = vals[i];
    vals[i] = run;
    run += h;
  }
  const int32_t total = part[src][1023];
  __syncthreads();
  return carry + total;
}

__global__ __launch_bounds__(1024) void key_scan_kernel(KeySort k) {
  __shared__ int32_t vals[kScanChunk];
  __shared__ int32_t part[2][1024];
  const int t = threadIdx.x;
  int32_t carry = 0;  // keys of the earlier passes
  for (int c0 = 0; c0 < k.n_keys; c0 += kScanChunk) {
    const int nc = min(kScanChunk, k.n_keys - c0);
    for (int i = t; i < nc; i += 1024) vals[i] = k.hist[c0 + i];
    __syncthreads();
    carry = scan_pass(vals, part, nc, carry, t);
    for (int i = t; i < nc; i += 1024) {
      k.cursor[c0 + i] = vals[i];
      k.hist[c0 + i] = 0;
    }
    __syncthreads();  // vals and part are rewritten by the next pass
  }
  if (t < 16 * k.segments) k.seg[t] = 0;
}

// Histogram and scan in one workgroup, for batches small enough that one
// workgroup reads every key quickly (nq <= kHistScanMaxQueries) and whose
// counts fit its LDS (n_keys <= kHistScanMaxKeys, dynamic LDS of n_keys
// words): the counts are LDS atomics, the global histogram is not touched
// (it stays zero), the scan is one pass, and a launch is saved.
constexpr int64_t kHistScanMaxQueries = 32768;
constexpr int kHistScanMaxKeys = 32768;  // 128 KB of counts
__global__ __launch_bounds__(1024) void key_hist_scan_kernel(KeySort k) {
  extern __shared__ int32_t vals[];  // n_keys counts
  __shared__ int32_t part[2][1024];
  const int t = threadIdx.x;
  for (int i = t; i < k.n_keys; i += 1024) vals[i] = 0;
  __syncthreads();
  // the leaves and beams of kU queries per thread are loaded (at clamped,
  // valid indices) before any is counted: one memory round trip per kU
  // queries instead of one per query
  constexpr int kU = 8;
  for (int64_t q0 = t; q0 < k.nq; q0 += 1024 * kU) {
    int lf[kU], bm[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      lf[u] = 0;
      bm[u] = 32;
    }
    if (k.leaf) {
#pragma unroll
      for (int u = 0; u < kU; ++u) lf[u] = k.leaf[min(q0 + 1024 * u, k.nq - 1)];
    }
    if (k.beam) {
#pragma unroll
      for (int u = 0; u < kU; ++u) bm[u] = k.beam[min(q0 + 1024 * u, k.nq - 1)];
    }
#pragma unroll
    for (int u = 0; u < kU; ++u)
      if (q0 + 1024 * u < k.nq) atomicAdd(&vals[key_from(k, lf[u], bm[u])], 1);
  }
  __syncthreads();
  scan_pass(vals, part, k.n_keys, 0, t);
  for (int i = t; i < k.n_keys; i += 1024) k.cursor[i] = vals[i];
  if (t < 16 * k.segments) k.seg[t] = 0;
}

__global__ void key_scatter_kernel(KeySort k) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < k.nq) k.order[atomicAdd(&k.cursor[sort_key(k, q)], 1)] = (int32_t)q;
}

// Launch attributes of the kernels above, set on the current device when a
// context is created (before any launch, so never inside a stream capture).
static hipError_t set_sort_kernel_attributes() {
  return hipFuncSetAttribute((const void*)key_hist_scan_kernel,
                             hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)(kHistScanMaxKeys * sizeof(int32_t)));
}

// cursors: classes in descending beam order
__global__ void batch_offsets_kernel(int64_t* counter) {
  int64_t off = 0;
  for (int cl = 4; cl >= 0; --cl) {
    counter[threadIdx.x + 8 + cl] = off;
    off += counter[threadIdx.x + 1 + cl];
  }
}

// scatter query ids by class; one atomic per wave and class, lanes take
// consecutive slots by their rank in the wave
__global__ void batch_order_kernel(const int32_t* __restrict__ beam, int64_t nq,
                                   int64_t* counter, int32_t* __restrict__ order) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int cl = q < nq ? (31 - __clz(beam[q])) - 5 : -1;
  const int lane = threadIdx.x & 63;
  for (int c = 0; c < 5; ++c) {
    const uint64_t m = __ballot(cl == c);
    if (!m) continue;
    const int leader = __ffsll((unsigned long long)m) - 1;
    unsigned long long base = 0;
    if (lane == leader)
      base = atomicAdd(reinterpret_cast<unsigned long long*>(counter + 8 + c),
                       (unsigned long long)__popcll(m));
    base = __shfl(base, leader, 64);
    if (cl == c) {
      const int rank = (int)__builtin_amdgcn_mbcnt_hi(
          (uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
      order[base + rank] = (int32_t)q;
    }
  }
}

// Batch preparation + persistent search, captured once per launch shape as a
// hipGraph and replayed (counters are cleared by a kernel, not a memset).
__global__ void zero_words_kernel(int32_t* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// The batch preparation kernels and the search, in stream order on c->stream
// (captured by launch_batched_graph, or launched directly).
static hipError_t enqueue_batched(rte_ctx* c, const rte::SearchArgs& a, const BatchPrep& bp,
                                  const KeySort* ks, int beam_max, int grid, int wpb) {
  const unsigned nb = (unsigned)((bp.nq + 255) / 256);
  if (bp.k_in && !ks) {  // uneven batch, ordered by beam class
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(1), 0, c->stream, a.work_counter);
    hipLaunchKernelGGL(batch_prep_kernel, dim3(nb), dim3(256), 0, c->stream, bp,
                       (int32_t*)a.k_per_q, (int32_t*)a.beam_per_q, a.work_counter);
    hipLaunchKernelGGL(batch_offsets_kernel, dim3(1), dim3(1), 0, c->stream, a.work_counter);
    hipLaunchKernelGGL(batch_order_kernel, dim3(nb), dim3(256), 0, c->stream, a.beam_per_q,
                       bp.nq, a.work_counter, (int32_t*)a.order);
  } else if (ks) {  // ordered by (beam class,) entry leaf
    if (bp.k_in)
      hipLaunchKernelGGL(batch_prep_kernel, dim3(nb), dim3(256), 0, c->stream, bp,
                         (int32_t*)a.k_per_q, (int32_t*)a.beam_per_q, (int64_t*)nullptr);
    if (ks->n_keys <= kHistScanMaxKeys && ks->nq <= kHistScanMaxQueries) {
      hipLaunchKernelGGL(key_hist_scan_kernel, dim3(1), dim3(1024),
                         (size_t)ks->n_keys * sizeof(int32_t), c->stream, *ks);
    } else {
      hipLaunchKernelGGL(key_hist_kernel, dim3(nb), dim3(256), 0, c->stream, *ks);
      hipLaunchKernelGGL(key_scan_kernel, dim3(1), dim3(1024), 0, c->stream, *ks);
    }
    hipLaunchKernelGGL(key_scatter_kernel, dim3(nb), dim3(256), 0, c->stream, *ks);
  }
  const hipError_t e1 = hipGetLastError();
  if (e1 != hipSuccess) return e1;
  return rte::launch_search(a, beam_max, grid, wpb, c->stream);
}

static int launch_batched_graph(rte_ctx* c, const rte::SearchArgs& a, const BatchPrep& bp,
                                const KeySort* ks, int beam_max, int grid, int wpb) {
  KeySort ksv{};
  if (ks) ksv = *ks;
  std::vector<unsigned char> key(sizeof(a) + sizeof(bp) + sizeof(ksv) + 5 * sizeof(int) +
                                 sizeof(void*));
  unsigned char* k = key.data();
  memcpy(k, &a, sizeof(a));
  k += sizeof(a);
  memcpy(k, &bp, sizeof(bp));
  k += sizeof(bp);
  memcpy(k, &ksv, sizeof(ksv));
  k += sizeof(ksv);
  const int extra[5] = {beam_max, grid, wpb, c->device, ks != nullptr};
  memcpy(k, extra, sizeof(extra));
  k += sizeof(extra);
  void* sp = (void*)c->stream;
  memcpy(k, &sp, sizeof(sp));
  if (!c->gexec || key != c->gkey) {
    if (c->gexec) {
      (void)hipGraphExecDestroy(c->gexec);
      c->gexec = nullptr;
    }
    hipGraph_t g = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed), "begin capture");
    const hipError_t e2 = enqueue_batched(c, a, bp, ks, beam_max, grid, wpb);
    const hipError_t e3 = hipStreamEndCapture(c->stream, &g);
    if (e2 != hipSuccess || e3 != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return hip_fail(c, e2 != hipSuccess ? e2 : e3, "capture batched");
    }
    const hipError_t e4 = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e4 != hipSuccess) {
      c->gexec = nullptr;
      return hip_fail(c, e4, "instantiate batched graph");
    }
    c->gkey = key;
  }
  HIPCHK(c, hipGraphLaunch(c->gexec, c->stream), "graph launch");
  return RTE_OK;
}

// Query ranges of RTE_ORDER_LEAF: one per XCD (workgroups are dealt to the
// eight XCDs of the part round-robin by blockIdx)
constexpr int kSegments = 8;

// Workgroup count for a grid code (rte_params.grid), from the occupancy of
// the kernel variant actually launched.
static int resolve_grid(rte_ctx* c, const rte::SearchArgs& a, int beam_max, int wpb, int code,
                        int64_t nq, bool batched, int* grid) {
  const int qpb = a.wpq > 1 ? 1 : wpb;  // queries per workgroup
  const int64_t per_query = std::min<int64_t>((nq + qpb - 1) / qpb, 1 << 20);
  if (code > 0) {
    *grid = code;
    return RTE_OK;
  }
  if (code == RTE_GRID_PER_QUERY || (code == RTE_GRID_DEFAULT && !batched)) {
    *grid = (int)per_query;
    return RTE_OK;
  }
  int occ = 0, cus = 0;
  HIPCHK(c, rte::search_occupancy(a, beam_max, wpb, &occ), "occupancy");
  HIPCHK(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device),
         "CU count");
  const int64_t resident = (int64_t)std::max(1, occ) * std::max(1, cus);
  int64_t g = resident;
  if (code == RTE_GRID_RESIDENT2) {
    g = 2 * resident;
  } else if (code == RTE_GRID_BALANCED) {
    // every round through the queue full: 10k queries on 4096 resident
    // units run 3334 units for 3 rounds instead of 4096 + 4096 + 1808
    const int64_t units = resident * qpb;
    const int64_t rounds = (nq + units - 1) / units;
    const int64_t need = (nq + rounds - 1) / rounds;
    g = (need + qpb - 1) / qpb;
  }
  *grid = (int)std::max<int64_t>(1, std::min<int64_t>(g, 1 << 20));
  return RTE_OK;
}

// Device buffers of a batched launch.
struct BatchBufs {
  int64_t* counter;  // 16: work counter (+ class counts and cursors)
  int64_t* seg;      // 16 * kSegments range counters
  int32_t* hist;     // n_keys
  int32_t* cursor;   // n_keys
  int32_t* k;        // nq effective k
  int32_t* beam;     // nq effective beams
  int32_t* order;    // nq query order
};

// Batched-launch buffers carved from one caller-provided device array
// (layout in ctx.h, BatchScratch), for calls that must not allocate.
static int scratch_bufs(rte_ctx* c, const rte::BatchScratch& sc, int64_t nq, int n_keys,
                        int where, BatchBufs* bb) {
  if (where != RTE_DEVICE)
    return fail(c, RTE_EINVAL, "a call without allocation takes device buffers only");
  if (nq > sc.max_queries)
    return fail(c, RTE_EINVAL, "%lld queries above the limit %lld of this call",
                (long long)nq, (long long)sc.max_queries);
  if (n_keys > sc.max_keys)
    return fail(c, RTE_EINVAL, "%d leaf keys above the limit %d of this call", n_keys,
                sc.max_keys);
  int32_t* base = sc.base;
  bb->counter = reinterpret_cast<int64_t*>(base);
  bb->seg = reinterpret_cast<int64_t*>(base + 2 * 16);
  bb->hist = base + rte::kScratchCounterWords;
  bb->cursor = bb->hist + sc.max_keys;
  bb->k = bb->cursor + sc.max_keys;
  bb->beam = bb->k + sc.max_queries;
  bb->order = bb->beam + sc.max_queries;
  return RTE_OK;
}

static int search_common(rte_ctx* c, const float* PQ, int64_t nq, int64_t q_stride,
                         const int32_t* leaf, const int32_t* k_per_q, const int32_t* beam_per_q,
                         const rte_params* p, int k_stride, int32_t* ids_out, float* dists_out,
                         int32_t* stats_out, int where, bool batched,
                         const rte::BatchScratch* na_scratch = nullptr) {
  const bool uneven = beam_per_q != nullptr;
  if (int r = check_params(c, p, uneven)) return r;
  if (!c->base) return fail(c, RTE_ESTATE, "rte_set_base not called");
  if (!c->graph) return fail(c, RTE_ESTATE, "rte_set_graph not called");
  if (c->graph_np != c->np)
    return fail(c, RTE_EINVAL, "graph has %lld rows, base %lld", (long long)c->graph_np,
                (long long)c->np);
  if (p->algo == 1 && !c->pids) return fail(c, RTE_ESTATE, "rte_set_leaves not called");
  if ((nq > 0 && (!PQ || !ids_out)) || nq < 0 || q_stride < c->d)
    return fail(c, RTE_EINVAL, "bad search args (q_stride %lld < d %d?)", (long long)q_stride, c->d);
  if (nq > 0 && p->algo == 1 && !leaf)
    return fail(c, RTE_EINVAL, "leaf array required for algo 1");
  if (nq == 0) return RTE_OK;
  if (int r = set_device(c)) return r;

  rte::SearchArgs a{};
  a.base = c->base;
  a.np = c->np;
  a.stride = c->stride;
  a.d = c->d;
  a.graph = c->graph;
  a.shift = c->shift;
  a.q_stride = q_stride;
  a.nq = nq;
  a.pids = c->pids;
  a.cap = c->cap;
  a.n_leaf = c->n_leaf;
  a.width = p->width;
  a.max_iter = p->max_iter;
  a.mode = p->mode;
  a.algo = p->algo == 1 ? 1 : 0;
  const int m = p->width << c->shift;
  int beam_max;
  std::vector<int32_t> hk, hb;
  if (!uneven) {
    // graph.cu:53: any n_candidates, rounded up to a power of two (beams
    // below 32 use the first min(beam, 32) bits of the expansion mask, P2)
    const int beam = pow2_ceil(std::max(p->beam, 1));
    a.beam = beam;
    a.k = p->k;
    a.k_stride = p->k;
    if (batched) {  // uniform batch: rows of k_stride >= k
      if (k_stride < p->k) return fail(c, RTE_EINVAL, "k_stride %d < k %d", k_stride, p->k);
      a.k_stride = k_stride;
    }
    beam_max = beam;
    int w, h;
    bloom_for(p, beam, &w, &h);
    a.bloom_words = w;
    a.bloom_hashes = h;
    a.bloom_words_max = w;
  } else {
    // per-query shapes are clamped on the device (batch_prep_kernel):
    // k into [1, k_stride], beam to pow2ceil(max(beam, k, 32)) <= the bound
    if (k_stride < 1) return fail(c, RTE_EINVAL, "k_stride < 1");
    beam_max = p->beam > 0 ? pow2_ceil(std::max(p->beam, 32)) : 512;
    if (beam_max > 512) return fail(c, RTE_EINVAL, "beam bound %d above 512", p->beam);
    const int kmax = std::min(k_stride, beam_max);
    a.k_stride = k_stride;
    a.beam = beam_max;
    a.k = kmax;
    if (p->mode == RTE_MODE_TUNED && p->bloom_words == 0) {
      a.bloom_auto = 1;
      int wmax = 0;
      for (int lg = 5; lg <= 9; ++lg) {
        int w, h;
        bloom_for(p, 1 << lg, &w, &h);  // a power of two below 2^15, 7 hashes
        a.auto_lgw |= (uint64_t)ilog2(w) << (4 * lg);
        wmax = std::max(wmax, w);
      }
      a.bloom_words_max = wmax;
    } else {
      int w, h;
      bloom_for(p, beam_max, &w, &h);
      a.bloom_words = w;
      a.bloom_hashes = h;
      a.bloom_words_max = w;
    }
  }
  if (!rte::search_supported(beam_max, m, c->d))
    return fail(c, RTE_EUNSUP, "no kernel for beam %d batch %d d %d", beam_max, m, c->d);

  const float* dq;
  const int32_t* dleaf = nullptr;
  int32_t* dids;
  float* dd = nullptr;
  int32_t* dst = nullptr;
  if (int r = stage_in(c, c->s_in, PQ, (size_t)nq * q_stride, where, &dq)) return r;
  if (p->algo == 1)
    if (int r = stage_in(c, c->s_aux, leaf, (size_t)nq, where, &dleaf)) return r;
  if (int r = stage_out(c, c->s_out, ids_out, (size_t)nq * a.k_stride, where, &dids)) return r;
  if (dists_out)
    if (int r = stage_out(c, c->s_aux2, dists_out, (size_t)nq * a.k_stride, where, &dd)) return r;
  if (stats_out)
    if (int r = stage_out(c, c->s_aux3, stats_out, (size_t)nq * 3, where, &dst)) return r;
  if (p->wpq != 0 && p->wpq != 1 && p->wpq != 2 && p->wpq != 4)
    return fail(c, RTE_EINVAL, "wpq %d must be 0, 1, 2 or 4", p->wpq);
  a.wpq = std::max(1, p->wpq);
  int wpb = p->wg_waves > 0 ? p->wg_waves : 4;
  if (wpb > 4) return fail(c, RTE_EINVAL, "wg_waves %d outside 1..4", p->wg_waves);
  if (a.wpq > 1) {  // the workgroup is the waves of one query
    if (p->wg_waves > 0 && p->wg_waves != a.wpq)
      return fail(c, RTE_EINVAL, "wg_waves %d must be 0 or wpq %d", p->wg_waves, a.wpq);
    wpb = a.wpq;
  }
  if (p->grid < RTE_GRID_PER_QUERY) return fail(c, RTE_EINVAL, "unknown grid code %d", p->grid);
  if (p->order != RTE_ORDER_GIVEN && p->order != RTE_ORDER_LEAF)
    return fail(c, RTE_EINVAL, "unknown order %d", p->order);
  if (p->queue != RTE_QUEUE_VGPR && p->queue != RTE_QUEUE_LDS)
    return fail(c, RTE_EINVAL, "unknown queue %d", p->queue);
  if (p->queue == RTE_QUEUE_LDS && a.wpq > 1)
    return fail(c, RTE_EINVAL, "the LDS queue runs one wave per query (wpq %d)", p->wpq);
  a.queue = p->queue;
  a.queries = dq;
  a.hits = dleaf;
  a.ids_out = dids;
  a.dists_out = dd;
  a.stats_out = dst;
  // the leaf order needs leaves; a batch of random entries keeps its order
  const bool by_leaf = batched && p->order == RTE_ORDER_LEAF && p->algo == 1 && c->n_leaf > 0;
  if (!batched) {
    int grid = 0;
    if (int r = resolve_grid(c, a, beam_max, wpb, p->grid, nq, false, &grid)) return r;
    HIPCHK(c, rte::launch_search(a, beam_max, grid, wpb, c->stream), "search launch");
  } else {
    const int n_keys = by_leaf ? c->n_leaf * (uneven ? 5 : 1) : 0;
    BatchBufs bb{};
    if (na_scratch) {  // the caller's no-allocation call: the static scratch
      if (int r = scratch_bufs(c, *na_scratch, nq, n_keys, where, &bb)) return r;
    } else {
      if (!c->b_counter.p) {
        HIPCHK(c, c->b_counter.reserve(16 * sizeof(int64_t)), "hipMalloc(counter)");
        HIPCHK(c, hipMemsetAsync(c->b_counter.p, 0, 16 * sizeof(int64_t), c->stream), "counter");
      }
      bb.counter = c->b_counter.as<int64_t>();
      if (uneven || by_leaf) {
        HIPCHK(c, c->b_order.reserve(nq * sizeof(int32_t)), "hipMalloc(order)");
        bb.order = c->b_order.as<int32_t>();
      }
      if (uneven) {
        HIPCHK(c, c->b_k.reserve(nq * sizeof(int32_t)), "hipMalloc(k_per_q)");
        HIPCHK(c, c->b_beam.reserve(nq * sizeof(int32_t)), "hipMalloc(beam_per_q)");
        bb.k = c->b_k.as<int32_t>();
        bb.beam = c->b_beam.as<int32_t>();
      }
      if (by_leaf) {
        HIPCHK(c, c->b_cursor.reserve((size_t)n_keys * sizeof(int32_t)), "hipMalloc(cursor)");
        if (c->b_hist.bytes < (size_t)n_keys * sizeof(int32_t)) {
          HIPCHK(c, c->b_hist.reserve((size_t)n_keys * sizeof(int32_t)), "hipMalloc(hist)");
          HIPCHK(c, hipMemsetAsync(c->b_hist.p, 0, c->b_hist.bytes, c->stream), "hist");
        }
        HIPCHK(c, c->b_seg.reserve(16 * kSegments * sizeof(int64_t)), "hipMalloc(ranges)");
        bb.hist = c->b_hist.as<int32_t>();
        bb.cursor = c->b_cursor.as<int32_t>();
        bb.seg = c->b_seg.as<int64_t>();
      }
    }
    a.work_counter = bb.counter;
    BatchPrep bp{nullptr, nullptr, nq, k_stride, beam_max};
    if (uneven) {
      // persistent launch: waves pull queries from a counter, longest beams
      // first (order built on the device), so mixed beams balance across CUs
      const int32_t* kin;
      const int32_t* bin;
      if (int r = stage_in(c, c->b_kin, k_per_q, (size_t)nq, where, &kin)) return r;
      if (int r = stage_in(c, c->b_bin, beam_per_q, (size_t)nq, where, &bin)) return r;
      a.k_per_q = bb.k;
      a.beam_per_q = bb.beam;
      a.order = bb.order;
      bp.k_in = kin;
      bp.beam_in = bin;
    }
    // a uniform batch in the given order is the search kernel alone; its
    // waves take queries from the self-clearing work counter
    KeySort ks{};
    if (by_leaf) {
      a.order = bb.order;
      a.work_counter = bb.seg;
      a.segments = kSegments;
      ks.leaf = dleaf;
      ks.beam = uneven ? a.beam_per_q : nullptr;
      ks.nq = nq;
      ks.n_leaf = c->n_leaf;
      ks.n_keys = n_keys;
      ks.hist = bb.hist;
      ks.cursor = bb.cursor;
      ks.order = bb.order;
      ks.seg = bb.seg;
      ks.segments = kSegments;
    }
    int grid = 0;
    if (int r = resolve_grid(c, a, beam_max, wpb, p->grid, nq, true, &grid)) return r;
    if (na_scratch) {
      // direct launches; what must start at zero is cleared by a kernel:
      // the counter block, the range counters and the key histogram
      const int64_t nz = rte::kScratchCounterWords + n_keys;
      hipLaunchKernelGGL(zero_words_kernel, dim3((unsigned)((nz + 255) / 256)), dim3(256), 0,
                         c->stream, na_scratch->base, nz);
      HIPCHK(c, enqueue_batched(c, a, bp, by_leaf ? &ks : nullptr, beam_max, grid, wpb),
             "batched launch");
    } else if (!uneven) {
      // a uniform batch (the search kernel, after the leaf sort when ordered
      // by leaf): direct launches.  Replaying the same kernels as a graph
      // left the GPU idle about 5.5 us before the graph's first kernel and
      // 8-13 us after its last one in kernel traces of the 10k-query step.
      HIPCHK(c, enqueue_batched(c, a, bp, by_leaf ? &ks : nullptr, beam_max, grid, wpb),
             "batched launch");
    } else if (int r = launch_batched_graph(c, a, bp, by_leaf ? &ks : nullptr, beam_max, grid,
                                            wpb)) {
      return r;
    }
  }
  if (int r = finish_out(c, ids_out, dids, (size_t)nq * a.k_stride, where)) return r;
  if (dists_out)
    if (int r = finish_out(c, dists_out, dd, (size_t)nq * a.k_stride, where)) return r;
  if (stats_out)
    if (int r = finish_out(c, stats_out, dst, (size_t)nq * 3, where)) return r;
  return RTE_OK;
}

int rte_search(rte_ctx* c, const float* PQ, int64_t nq, int64_t q_stride, const int32_t* leaf,
               const rte_params* p, int32_t* ids_out, float* dists_out, int32_t* stats_out,
               int where) {
  if (!c) return RTE_EINVAL;
  return search_common(c, PQ, nq, q_stride, leaf, nullptr, nullptr, p, p ? p->k : 0, ids_out,
                       dists_out, stats_out, where, false);
}

int rte_search_batched(rte_ctx* c, const float* PQ, int64_t nq, int64_t q_stride,
                       const int32_t* leaf, const int32_t* k_per_q, const int32_t* beam_per_q,
                       const rte_params* p, int k_stride, int32_t* ids_out, float* dists_out,
                       int where) {
  if (!c) return RTE_EINVAL;
  if (!k_per_q != !beam_per_q)
    return fail(c, RTE_EINVAL, "k_per_q and beam_per_q: both or neither");
  return search_common(c, PQ, nq, q_stride, leaf, k_per_q, beam_per_q, p, k_stride, ids_out,
                       dists_out, nullptr, where, true);
}

}  // extern "C"

int rte_search_batched_in(rte_ctx* c, const float* PQ, int64_t nq, int64_t q_stride,
                          const int32_t* leaf, const int32_t* k_per_q, const int32_t* beam_per_q,
                          const rte_params* p, int k_stride, int32_t* ids_out, float* dists_out,
                          const rte::BatchScratch& sc) {
  if (!c) return RTE_EINVAL;
  if (!k_per_q != !beam_per_q)
    return fail(c, RTE_EINVAL, "k_per_q and beam_per_q: both or neither");
  return search_common(c, PQ, nq, q_stride, leaf, k_per_q, beam_per_q, p, k_stride, ids_out,
                       dists_out, nullptr, RTE_DEVICE, true, &sc);
}

extern "C" {

int rte_autotune_batched(rte_ctx* c, const float* PQ, int64_t nq, int64_t q_stride,
                         const int32_t* leaf, const int32_t* k_per_q,
                         const int32_t* beam_per_q, const rte_params* p, int k_stride,
                         int32_t* ids_out, float* dists_out, const int* bloom_words,
                         int n_bloom, rte_params* best, float* best_ms) {
  if (!c) return RTE_EINVAL;
  if (!p || !best || !best_ms || !bloom_words || n_bloom < 1)
    return fail(c, RTE_EINVAL, "bad autotune arguments");
  if (int r = set_device(c)) return r;
  // candidates: waves per query x queries per workgroup x beam storage x
  // grid code x query order x filter size; every grid code is resolved by
  // the call itself from the occupancy of the kernel it launches
  std::vector<rte_params> cands;
  const int orders[2] = {RTE_ORDER_GIVEN, RTE_ORDER_LEAF};
  for (int bi = 0; bi < n_bloom; ++bi)
    for (int wpq : {1, 2, 4})
      for (int wpb : {1, 2, 4})
        for (int qu : {RTE_QUEUE_VGPR, RTE_QUEUE_LDS}) {
          if (wpq > 1 && (wpb != wpq || qu == RTE_QUEUE_LDS)) continue;
          for (int g : {RTE_GRID_PER_QUERY, RTE_GRID_DEFAULT, RTE_GRID_RESIDENT2,
                        RTE_GRID_BALANCED})
            for (int o = 0; o < (p->algo == 1 ? 2 : 1); ++o) {
              rte_params t = *p;
              t.bloom_words = bloom_words[bi];
              t.wpq = wpq;
              t.wg_waves = wpb;
              t.queue = qu;
              t.grid = g;
              t.order = orders[o];
              cands.push_back(t);
            }
        }
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0), "event");
  HIPCHK(c, hipEventCreate(&e1), "event");
  // one timed launch of shape t (ms), or a negative value when it failed
  auto time_one = [&](const rte_params& t, int* rc) -> float {
    (void)hipEventRecord(e0, c->stream);
    *rc = rte_search_batched(c, PQ, nq, q_stride, leaf, k_per_q, beam_per_q, &t, k_stride,
                             ids_out, dists_out, RTE_DEVICE);
    (void)hipEventRecord(e1, c->stream);
    if (hipEventSynchronize(e1) != hipSuccess && *rc == RTE_OK)
      *rc = fail(c, RTE_EHIP, "autotune: event wait failed");
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    return *rc == RTE_OK ? ms : -1.f;
  };
  int first_rc = RTE_OK;
  std::string first_msg;
  std::vector<std::pair<float, size_t>> timed;  // (mean ms, candidate index)
  for (size_t ci = 0; ci < cands.size(); ++ci) {
    const rte_params& t = cands[ci];
    // the first run sets the launch up; then three timed runs
    int rc = rte_search_batched(c, PQ, nq, q_stride, leaf, k_per_q, beam_per_q, &t, k_stride,
                                ids_out, dists_out, RTE_DEVICE);
    float ms_sum = 0.f;
    for (int rep = 0; rep < 3 && rc == RTE_OK; ++rep) ms_sum += time_one(t, &rc);
    if (rc != RTE_OK) {  // this shape failed: remember why, try the others
      if (first_rc == RTE_OK) {
        first_rc = rc;
        first_msg = c->err;
      }
      continue;
    }
    timed.emplace_back(ms_sum / 3.f, ci);
  }
  if (timed.empty()) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->err = first_msg;
    return first_rc;
  }
  // Second stage: shapes within kTuneMargin of the fastest mean differ by
  // about as much as one launch varies from the next, so the finalists (at
  // most kTuneFinal) are timed again in kTuneRounds interleaved rounds --
  // each round runs every finalist once, so a drift of the clock or of the
  // memory system reaches all of them alike -- and the shape with the lowest
  // median wins.
  constexpr float kTuneMargin = 1.05f;
  constexpr size_t kTuneFinal = 6;
  constexpr int kTuneRounds = 9;
  std::sort(timed.begin(), timed.end());
  size_t nf = 1;
  while (nf < timed.size() && nf < kTuneFinal && timed[nf].first <= timed[0].first * kTuneMargin)
    ++nf;
  size_t pick = timed[0].second;
  float best_t = timed[0].first;
  if (nf > 1) {
    std::vector<std::vector<float>> ms(nf);
    std::vector<bool> alive(nf, true);
    for (int round = 0; round < kTuneRounds; ++round)
      for (size_t f = 0; f < nf; ++f) {
        if (!alive[f]) continue;
        int rc = RTE_OK;
        const float t = time_one(cands[timed[f].second], &rc);
        if (rc != RTE_OK) {
          alive[f] = false;  // it ran in the first stage; drop it if it fails now
          continue;
        }
        ms[f].push_back(t);
      }
    float best_med = 1e30f;
    for (size_t f = 0; f < nf; ++f) {
      if (!alive[f] || ms[f].empty()) continue;
      std::vector<float> v = ms[f];
      std::nth_element(v.begin(), v.begin() + v.size() / 2, v.end());
      const float med = v[v.size() / 2];
      if (med < best_med) {
        best_med = med;
        pick = timed[f].second;
      }
    }
    if (best_med < 1e30f) best_t = best_med;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *best = cands[pick];
  // leave the outputs of the chosen configuration
  if (int r = rte_search_batched(c, PQ, nq, q_stride, leaf, k_per_q, beam_per_q, best, k_stride,
                                 ids_out, dists_out, RTE_DEVICE))
    return r;
  *best_ms = best_t;
  return RTE_OK;
}

int rte_query(rte_ctx* c, const float* Q, int64_t nq, const rte_params* p, float* PQ_scratch,
              int32_t* leaf_scratch, int32_t* ids_out, float* dists_out) {
  if (!c) return RTE_EINVAL;
  if (int r = rte_project(c, Q, nq, PQ_scratch, RTE_DEVICE)) return r;
  if (p && p->algo == 1)
    if (int r = rte_select_entry(c, PQ_scratch, nq, c->pca_out, leaf_scratch, RTE_DEVICE)) return r;
  return rte_search(c, PQ_scratch, nq, c->pca_out, leaf_scratch, p, ids_out, dists_out, nullptr,
                    RTE_DEVICE);
}

void rte_hash_parameter(int nc, int* hash_len, int* bit, int* hash) {
  hash_parameter(nc, hash_len, bit, hash);
}

int rte_search_supported(int beam, int batch, int d) {
  return rte::search_supported(pow2_ceil(std::max(beam, 1)), batch, d);
}

const char* rte_version(void) { return "rt_entry_amd 0.1 gfx950"; }

}  // extern "C"
