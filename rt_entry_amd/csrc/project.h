// FP32 MFMA projection GEMM (project.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rte {

// out[nq x d_out] = Q[nq x d] R[d x d_out] - bias[d_out]; R is row-major
// with row stride d_out.
hipError_t launch_project(const float* Q, int64_t nq, int d, const float* R, int d_out,
                          const float* bias, float* out, hipStream_t st);
hipError_t launch_mu_r(const float* mu, const float* R, int d, int d_out, float* bias,
                       hipStream_t st);

}  // namespace rte
