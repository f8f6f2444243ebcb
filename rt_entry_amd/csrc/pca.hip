// PCA fit on the GPU: FP64 mean and covariance of the base, symmetric
// eigen-decomposition on the host (Householder + QL, FP64), eigenvectors sorted
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

// Symmetric eigen-decomposition (FP64): the published EISPACK method,
// tred2 (Householder reduction to tridiagonal form, accumulating the
// transformations) followed by tql2 (QL iterations with implicit shifts on
// the tridiagonal matrix), in the widely printed tred2/tqli formulation
// (Martin, Reinsch and Wilkinson, Numer. Math. 11 and 12, 1968).  The
// reference's pca.cpp uses Eigen's SelfAdjointEigenSolver, a different
// formulation (tridiagonalisation then implicit symmetric QR steps with a
// Wilkinson shift); both give the eigenpairs to FP64 rounding, which the
// tests check against LAPACK.  Every inner loop walks rows: about 1.3 s at
// d = 960 and a few ms at d = 128.
//
// Householder step: A (d x d, row-major, only its lower triangle used) is
// destroyed; on return A holds Q with A_in = Q T Q^T, diag the diagonal of T
// and off[i] = T[i][i-1] (off[0] = 0).
static void householder_tridiag(std::vector<double>& A, int d, std::vector<double>& diag,
                                std::vector<double>& off) {
  diag.assign(d, 0.0);
  off.assign(d, 0.0);
  auto a = [&](int i, int j) -> double& { return A[(size_t)i * d + j]; };
  for (int i = d - 1; i > 0; --i) {
    const int l = i - 1;
    double h = 0.0;
    if (l > 0) {
      double scale = 0.0;
      for (int k = 0; k <= l; ++k) scale += fabs(a(i, k));
      if (scale == 0.0) {
        off[i] = a(i, l);
      } else {
        for (int k = 0; k <= l; ++k) {
          a(i, k) /= scale;
          h += a(i, k) * a(i, k);
        }
        const double f = a(i, l);
        const double g = f >= 0.0 ? -sqrt(h) : sqrt(h);
        off[i] = scale * g;
        h -= f * g;
        a(i, l) = f - g;
        // p = A[0..l][0..l] u (u = row i), from the lower triangle row by row
        const double* u = &A[(size_t)i * d];
        for (int j = 0; j <= l; ++j) off[j] = 0.0;
        for (int j = 0; j <= l; ++j) {
          const double* aj = &A[(size_t)j * d];
          double pj = aj[j] * u[j];
          for (int k = 0; k < j; ++k) {
            pj += aj[k] * u[k];
            off[k] += aj[k] * u[j];
          }
          off[j] += pj;
        }
        double ff = 0.0;
        for (int j = 0; j <= l; ++j) {
          a(j, i) = u[j] / h;
          off[j] /= h;
          ff += off[j] * u[j];
        }
        const double hh = ff / (h + h);
        for (int j = 0; j <= l; ++j) {  // A -= u q^T + q u^T on the lower triangle
          const double fj = a(i, j);
          const double gj = off[j] - hh * fj;
          off[j] = gj;
          for (int k = 0; k <= j; ++k) a(j, k) -= fj * off[k] + gj * a(i, k);
        }
      }
    } else {
      off[i] = a(i, l);
    }
    diag[i] = h;
  }
  diag[0] = 0.0;
  off[0] = 0.0;
  // accumulate the reflections into A, which becomes Q
  std::vector<double> gv(d);
  for (int i = 0; i < d; ++i) {
    const int l = i - 1;
    if (diag[i] != 0.0) {
      // g = a(i, 0..l) A[0..l][0..l], then A[k][0..l] -= a(k, i) g, row by row
      for (int j = 0; j <= l; ++j) gv[j] = 0.0;
      for (int k = 0; k <= l; ++k) {
        const double aik = a(i, k);
        const double* ak = &A[(size_t)k * d];
        for (int j = 0; j <= l; ++j) gv[j] += aik * ak[j];
      }
      for (int k = 0; k <= l; ++k) {
        const double aki = a(k, i);
        double* ak = &A[(size_t)k * d];
        for (int j = 0; j <= l; ++j) ak[j] -= aki * gv[j];
      }
    }
    diag[i] = a(i, i);
    a(i, i) = 1.0;
    for (int j = 0; j <= l; ++j) a(j, i) = a(i, j) = 0.0;
  }
}

// QL with implicit shifts on the tridiagonal (diag, off), rotating the rows
// of Zt (row i = vector i, Q^T on entry): on return diag holds the
// eigenvalues and row i of Zt the eigenvector of diag[i].  False if an
// eigenvalue needed more than 60 iterations.
static bool tridiag_ql(std::vector<double>& diag, std::vector<double>& off,
                       std::vector<double>& Zt, int d) {
  for (int i = 1; i < d; ++i) off[i - 1] = off[i];
  off[d - 1] = 0.0;
  for (int l = 0; l < d; ++l) {
    int iter = 0;
    int m;
    do {
      for (m = l; m < d - 1; ++m) {  // first negligible off-diagonal element
        const double dd = fabs(diag[m]) + fabs(diag[m + 1]);
        if (fabs(off[m]) <= 1e-15 * dd) break;
      }
      if (m != l) {
        if (iter++ == 60) return false;
        double g = (diag[l + 1] - diag[l]) / (2.0 * off[l]);
        double r = hypot(g, 1.0);
        g = diag[m] - diag[l] + off[l] / (g + (g >= 0.0 ? fabs(r) : -fabs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        for (i = m - 1; i >= l; --i) {
          double f = s * off[i];
          const double b = c * off[i];
          r = hypot(f, g);
          off[i + 1] = r;
          if (r == 0.0) {  // underflow: deflate and restart
            diag[i + 1] -= p;
            off[m] = 0.0;
            break;
          }
          s = f / r;
          c = g / r;
          g = diag[i + 1] - p;
          r = (diag[i] - g) * s + 2.0 * c * b;
          p = s * r;
          diag[i + 1] = g + p;
          g = c * r - b;
          double* zi = &Zt[(size_t)i * d];
          double* zj = zi + d;
          for (int k = 0; k < d; ++k) {
            f = zj[k];
            zj[k] = s * zi[k] + c * f;
            zi[k] = c * zi[k] - s * f;
          }
        }
        if (r == 0.0 && i >= l) continue;
        diag[l] -= p;
        off[l] = g;
        off[m] = 0.0;
      }
    } while (m != l);
  }
  return true;
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
  std::vector<double> w, off;
  rte::householder_tridiag(C, d, w, off);
  std::vector<double> Zt(dd);  // Q^T: row i = column i of Q
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j) Zt[(size_t)j * d + i] = C[(size_t)i * d + j];
  if (!rte::tridiag_ql(w, off, Zt, d))
    return rte_fail(c, RTE_EBUILD, "pca: eigensolver did not converge");
  std::vector<int> order(d);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return w[a] > w[b]; });
  for (int j = 0; j < d; ++j) {
    const double* v = &Zt[(size_t)order[j] * d];
    int arg = 0;  // deterministic sign: largest-magnitude component positive
    for (int i = 1; i < d; ++i)
      if (fabs(v[i]) > fabs(v[arg])) arg = i;
    const double sg = v[arg] < 0 ? -1.0 : 1.0;
    for (int i = 0; i < d; ++i) R_out[(size_t)i * d + j] = (float)(sg * v[i]);
    if (eig_out) eig_out[j] = w[order[j]];
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
