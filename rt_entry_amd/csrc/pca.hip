// PCA fit on the GPU: FP64 mean and covariance of the base, symmetric
// eigen-decomposition on the host (cyclic Jacobi, FP64), eigenvectors sorted
// by descending eigenvalue with a deterministic sign (largest-magnitude
// component positive).  Replaces the reference's Eigen-based PCA::PCA
// (pca.cpp:4-58: mean, covariance / (n - 1), SelfAdjointEigenSolver, columns
// reordered by descending eigenvalue), which holds the whole base as doubles
// on the host (pca.h:15-21).
//
// Two passes over the base.  Column sums: rows split over workgroups, the
// per-workgroup partials summed in a fixed order.  Covariance: the d x d
// result is cut into 128 x 128 tiles; a workgroup accumulates, for one tile
// on or above the diagonal and one range of rows, sum (x - mu)_i (x - mu)_j
// with one 8 x 8 sub-tile per thread, the centred rows staged through LDS 16
// at a time; the per-row-range partials are again summed in a fixed order
// and the lower triangle is the mirror image (the products are the same
// numbers in the same order, so C is exactly symmetric).  Everything is
// deterministic.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "ctx.h"

namespace rte {

constexpr int kPcaTile = 128;
constexpr int kPcaRowsLds = 16;
constexpr int kPcaMaxD = 4096;

__global__ __launch_bounds__(256) void col_sum_kernel(const float* __restrict__ X, int64_t n,
                                                      int d, int64_t stride, int64_t rpb,
                                                      double* __restrict__ partial) {
  const int64_t r0 = (int64_t)blockIdx.x * rpb;
  const int64_t r1 = r0 + rpb < n ? r0 + rpb : n;
  for (int c = threadIdx.x; c < d; c += blockDim.x) {
    double s = 0.0;
    for (int64_t r = r0; r < r1; ++r) s += (double)X[r * stride + c];
    partial[(int64_t)blockIdx.x * d + c] = s;
  }
}

__global__ void col_mean_kernel(const double* __restrict__ partial, int nb, int d, int64_t n,
                                double* __restrict__ mean) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d) return;
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s += partial[(int64_t)b * d + c];
  mean[c] = s / (double)n;
}

// grid (row ranges, tile pairs); tile pair p -> (bi, bj), bi <= bj, row-major
// over the upper triangle of the nt x nt tile grid
__global__ __launch_bounds__(256) void cov_partial_kernel(const float* __restrict__ X, int64_t n,
                                                          int d, int64_t stride,
                                                          const double* __restrict__ mean,
                                                          int64_t rpb, int nt,
                                                          double* __restrict__ partial) {
  __shared__ double ra[kPcaRowsLds][kPcaTile];
  __shared__ double rb[kPcaRowsLds][kPcaTile];
  int p = blockIdx.y, bi = 0;
  while (p >= nt - bi) {
    p -= nt - bi;
    ++bi;
  }
  const int bj = bi + p;
  const int c0 = bi * kPcaTile, c1 = bj * kPcaTile;
  const int t = threadIdx.x;
  const int ti = t >> 4, tj = t & 15;  // 8 x 8 sub-tile (ti, tj) of the 16 x 16 grid
  double acc[8][8];
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[i][j] = 0.0;
  const int64_t r0 = (int64_t)blockIdx.x * rpb;
  const int64_t r1 = r0 + rpb < n ? r0 + rpb : n;
  for (int64_t rs = r0; rs < r1; rs += kPcaRowsLds) {
    const int nr = (int)(r1 - rs < kPcaRowsLds ? r1 - rs : kPcaRowsLds);
    for (int e = t; e < kPcaRowsLds * kPcaTile; e += 256) {
      const int rr = e / kPcaTile, c = e - rr * kPcaTile;
      const bool ok = rr < nr;
      ra[rr][c] = ok && c0 + c < d ? (double)X[(rs + rr) * stride + c0 + c] - mean[c0 + c] : 0.0;
      rb[rr][c] = ok && c1 + c < d ? (double)X[(rs + rr) * stride + c1 + c] - mean[c1 + c] : 0.0;
    }
    __syncthreads();
    for (int rr = 0; rr < nr; ++rr) {
      double a[8], b[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        a[i] = ra[rr][8 * ti + i];
        b[i] = rb[rr][8 * tj + i];
      }
#pragma unroll
      for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[i][j] = __builtin_fma(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
  double* out = partial + (int64_t)blockIdx.x * d * d;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int gi = c0 + 8 * ti + i;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int gj = c1 + 8 * tj + j;
      if (gi < d && gj < d) out[(int64_t)gi * d + gj] = acc[i][j];
    }
  }
}

__global__ void cov_reduce_kernel(const double* __restrict__ partial, int nb, int d, int64_t n,
                                  double* __restrict__ C) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)d * d) return;
  const int i = (int)(e / d), j = (int)(e - (int64_t)i * d);
  // tiles below the diagonal were not computed: read the mirror element
  const int64_t src = i / kPcaTile > j / kPcaTile ? (int64_t)j * d + i : e;
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s += partial[(int64_t)b * d * d + src];
  C[e] = s / (double)(n > 1 ? n - 1 : 1);  // pca.cpp:8-9
}

// out[i][c] = sum_j ((double)X[i][j] - (double)mu[j]) * (double)R[j][c],
// j ascending, multiply then add (the library is built without contraction),
// stored FP32 -- the FP64 rotation of pca.cpp:122-139.  A workgroup computes
// 16 rows x 128 columns (8 rows per thread); the centred rows are staged
// through LDS 128 columns of X at a time.
__global__ __launch_bounds__(256) void pca_rotate_kernel(const float* __restrict__ X, int64_t n,
                                                         int d, int64_t stride,
                                                         const float* __restrict__ R,
                                                         const float* __restrict__ mu, int d_out,
                                                         float* __restrict__ out) {
  __shared__ double xs[kPcaRowsLds][kPcaTile];
  const int t = threadIdx.x;
  const int c = blockIdx.y * kPcaTile + (t & 127);
  const int half = t >> 7;
  const int64_t row0 = (int64_t)blockIdx.x * kPcaRowsLds;
  double acc[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) acc[i] = 0.0;
  for (int j0 = 0; j0 < d; j0 += kPcaTile) {
    for (int e = t; e < kPcaRowsLds * kPcaTile; e += 256) {
      const int rr = e / kPcaTile, j = j0 + (e - rr * kPcaTile);
      xs[rr][e - rr * kPcaTile] =
          row0 + rr < n && j < d ? (double)X[(row0 + rr) * stride + j] - (double)mu[j] : 0.0;
    }
    __syncthreads();
    const int jn = d - j0 < kPcaTile ? d - j0 : kPcaTile;
    if (c < d_out) {
      for (int j = 0; j < jn; ++j) {
        const double r = (double)R[(int64_t)(j0 + j) * d_out + c];
#pragma unroll
        for (int i = 0; i < 8; ++i) acc[i] = acc[i] + xs[8 * half + i][j] * r;
      }
    }
    __syncthreads();
  }
  if (c < d_out) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int64_t row = row0 + 8 * half + i;
      if (row < n) out[row * d_out + c] = (float)acc[i];
    }
  }
}

// Cyclic Jacobi eigen-decomposition of a symmetric d x d matrix (FP64):
// A is destroyed, V gets the eigenvectors as columns, w the eigenvalues.
// From the fifth sweep on, an off-diagonal element too small to change
// either diagonal element is set to zero without a rotation, so the sweeps
// end with an exactly diagonal matrix.
static void jacobi_eigen(std::vector<double>& A, int d, std::vector<double>& V,
                         std::vector<double>& w) {
  V.assign((size_t)d * d, 0.0);
  for (int i = 0; i < d; ++i) V[(size_t)i * d + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int i = 0; i < d; ++i)
      for (int j = i + 1; j < d; ++j) off += fabs(A[(size_t)i * d + j]);
    if (off == 0.0) break;
    for (int p = 0; p < d - 1; ++p)
      for (int q = p + 1; q < d; ++q) {
        const double apq = A[(size_t)p * d + q];
        const double app = A[(size_t)p * d + p], aqq = A[(size_t)q * d + q];
        const double g = 100.0 * fabs(apq);
        if (sweep >= 4 && fabs(app) + g == fabs(app) && fabs(aqq) + g == fabs(aqq)) {
          A[(size_t)p * d + q] = A[(size_t)q * d + p] = 0.0;
          continue;
        }
        if (apq == 0.0) continue;
        const double theta = (aqq - app) / (2.0 * apq);
        const double tt = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(tt * tt + 1.0), s = tt * c;
        for (int k = 0; k < d; ++k) {  // columns p, q
          const double akp = A[(size_t)k * d + p], akq = A[(size_t)k * d + q];
          A[(size_t)k * d + p] = c * akp - s * akq;
          A[(size_t)k * d + q] = s * akp + c * akq;
        }
        for (int k = 0; k < d; ++k) {  // rows p, q
          const double apk = A[(size_t)p * d + k], aqk = A[(size_t)q * d + k];
          A[(size_t)p * d + k] = c * apk - s * aqk;
          A[(size_t)q * d + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < d; ++k) {
          const double vkp = V[(size_t)k * d + p], vkq = V[(size_t)k * d + q];
          V[(size_t)k * d + p] = c * vkp - s * vkq;
          V[(size_t)k * d + q] = s * vkp + c * vkq;
        }
      }
  }
  w.resize(d);
  for (int i = 0; i < d; ++i) w[i] = A[(size_t)i * d + i];
}

}  // namespace rte

extern "C" {

int rte_pca_fit(rte_ctx* c, const float* X, int64_t n, int d, int64_t stride, float* R_out,
                float* mu_out, double* eig_out, int where) {
  if (!c) return RTE_EINVAL;
  if (!X || !R_out || !mu_out || n < 2 || d < 1 || stride < d)
    return rte_fail(c, RTE_EINVAL, "bad pca arguments");
  if (d > rte::kPcaMaxD) return rte_fail(c, RTE_EUNSUP, "pca for d <= %d (d = %d)", rte::kPcaMaxD, d);
  if (hipSetDevice(c->device) != hipSuccess) return rte_fail(c, RTE_EHIP, "hipSetDevice");
  hipStream_t st = c->stream;
  auto ck = [&](hipError_t e, const char* what) {
    return e == hipSuccess ? RTE_OK : rte_fail(c, RTE_EHIP, "%s: %s", what, hipGetErrorString(e));
  };
  DevBuf xin, part, mean, cov;
  const float* dx = X;
  if (where == RTE_HOST) {
    const size_t bytes = ((size_t)(n - 1) * stride + d) * sizeof(float);
    if (int r = ck(xin.reserve(bytes), "hipMalloc(pca input)")) return r;
    if (int r = ck(hipMemcpyAsync(xin.p, X, bytes, hipMemcpyHostToDevice, st), "copy")) return r;
    dx = xin.as<float>();
  }
  int cus = 0;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device);
  cus = std::max(cus, 1);
  const int nt = (d + rte::kPcaTile - 1) / rte::kPcaTile;
  const int ntp = nt * (nt + 1) / 2;
  // row ranges: enough workgroups to fill the chip twice, >= 1024 rows each
  const int64_t by_rows = std::max<int64_t>(1, (n + 1023) / 1024);
  const int nb_sum = (int)std::min<int64_t>(2 * cus, by_rows);
  const int nb = (int)std::min<int64_t>(std::max(1, (2 * cus + ntp - 1) / ntp), by_rows);
  const int64_t rpb_sum = (n + nb_sum - 1) / nb_sum, rpb = (n + nb - 1) / nb;
  const size_t dd = (size_t)d * d;
  if (int r = ck(part.reserve(std::max((size_t)nb * dd, (size_t)nb_sum * d) * sizeof(double)),
                 "hipMalloc(pca partials)"))
    return r;
  if (int r = ck(mean.reserve((size_t)d * sizeof(double)), "hipMalloc")) return r;
  if (int r = ck(cov.reserve(dd * sizeof(double)), "hipMalloc")) return r;
  hipLaunchKernelGGL(rte::col_sum_kernel, dim3(nb_sum), dim3(256), 0, st, dx, n, d, stride,
                     rpb_sum, part.as<double>());
  hipLaunchKernelGGL(rte::col_mean_kernel, dim3((d + 127) / 128), dim3(128), 0, st,
                     part.as<double>(), nb_sum, d, n, mean.as<double>());
  hipLaunchKernelGGL(rte::cov_partial_kernel, dim3(nb, ntp), dim3(256), 0, st, dx, n, d, stride,
                     mean.as<double>(), rpb, nt, part.as<double>());
  hipLaunchKernelGGL(rte::cov_reduce_kernel, dim3((unsigned)((dd + 255) / 256)), dim3(256), 0, st,
                     part.as<double>(), nb, d, n, cov.as<double>());
  if (int r = ck(hipGetLastError(), "pca kernels")) return r;
  std::vector<double> C(dd), mu(d);
  if (int r = ck(hipMemcpyAsync(C.data(), cov.p, dd * 8, hipMemcpyDeviceToHost, st), "copy"))
    return r;
  if (int r = ck(hipMemcpyAsync(mu.data(), mean.p, (size_t)d * 8, hipMemcpyDeviceToHost, st),
                 "copy"))
    return r;
  if (int r = ck(hipStreamSynchronize(st), "sync")) return r;
  std::vector<double> V, w;
  rte::jacobi_eigen(C, d, V, w);
  std::vector<int> order(d);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return w[a] > w[b]; });
  for (int j = 0; j < d; ++j) {
    const int src = order[j];
    int arg = 0;  // deterministic sign: largest-magnitude component positive
    for (int i = 1; i < d; ++i)
      if (fabs(V[(size_t)i * d + src]) > fabs(V[(size_t)arg * d + src])) arg = i;
    const double sg = V[(size_t)arg * d + src] < 0 ? -1.0 : 1.0;
    for (int i = 0; i < d; ++i) R_out[(size_t)i * d + j] = (float)(sg * V[(size_t)i * d + src]);
    if (eig_out) eig_out[j] = w[src];
  }
  for (int j = 0; j < d; ++j) mu_out[j] = (float)mu[j];
  return RTE_OK;
}

int rte_pca_rotate(rte_ctx* c, const float* X, int64_t n, int d, int64_t stride, const float* R,
                   const float* mu, int d_out, float* out, int where) {
  if (!c) return RTE_EINVAL;
  if (!X || !R || !mu || !out || n < 0 || d < 1 || d_out < 1 || d_out > d || stride < d)
    return rte_fail(c, RTE_EINVAL, "bad pca_rotate arguments");
  if (n == 0) return RTE_OK;
  if (hipSetDevice(c->device) != hipSuccess) return rte_fail(c, RTE_EHIP, "hipSetDevice");
  hipStream_t st = c->stream;
  auto ck = [&](hipError_t e, const char* what) {
    return e == hipSuccess ? RTE_OK : rte_fail(c, RTE_EHIP, "%s: %s", what, hipGetErrorString(e));
  };
  if ((n + rte::kPcaRowsLds - 1) / rte::kPcaRowsLds > 0x7fffffff)
    return rte_fail(c, RTE_EUNSUP, "pca_rotate: too many rows");
  DevBuf xin, rm, ob;
  const float* dx = X;
  float* dout = out;
  const size_t xbytes = ((size_t)(n - 1) * stride + d) * sizeof(float);
  const size_t obytes = (size_t)n * d_out * sizeof(float);
  if (where == RTE_HOST) {
    if (int r = ck(xin.reserve(xbytes), "hipMalloc(pca_rotate input)")) return r;
    if (int r = ck(ob.reserve(obytes), "hipMalloc(pca_rotate output)")) return r;
    if (int r = ck(hipMemcpyAsync(xin.p, X, xbytes, hipMemcpyHostToDevice, st), "copy")) return r;
    dx = xin.as<float>();
    dout = ob.as<float>();
  }
  // R (d x d_out) and mu (d) are host arrays
  if (int r = ck(rm.reserve(((size_t)d * d_out + d) * sizeof(float)), "hipMalloc")) return r;
  float* dR = rm.as<float>();
  float* dmu = dR + (size_t)d * d_out;
  if (int r = ck(hipMemcpyAsync(dR, R, (size_t)d * d_out * 4, hipMemcpyHostToDevice, st), "copy"))
    return r;
  if (int r = ck(hipMemcpyAsync(dmu, mu, (size_t)d * 4, hipMemcpyHostToDevice, st), "copy"))
    return r;
  dim3 grid((unsigned)((n + rte::kPcaRowsLds - 1) / rte::kPcaRowsLds),
            (unsigned)((d_out + rte::kPcaTile - 1) / rte::kPcaTile));
  hipLaunchKernelGGL(rte::pca_rotate_kernel, grid, dim3(256), 0, st, dx, n, d, stride, dR, dmu,
                     d_out, dout);
  if (int r = ck(hipGetLastError(), "pca_rotate kernel")) return r;
  if (where == RTE_HOST) {
    if (int r = ck(hipMemcpyAsync(out, dout, obytes, hipMemcpyDeviceToHost, st), "copy")) return r;
  }
  return ck(hipStreamSynchronize(st), "sync");
}

}  // extern "C"
