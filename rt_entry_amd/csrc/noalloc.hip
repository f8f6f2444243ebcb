// rte_search_batched_noalloc: the batched search for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving
// loops).  Every launch buffer of the batch lives in one static device
// array per GPU; nothing here calls hipMalloc/hipFree or captures a graph.
#include <hip/hip_runtime.h>

#include <mutex>

#include "ctx.h"
#include "rt_entry.h"

namespace {

constexpr int64_t kMaxQueries = RTE_NOALLOC_MAX_QUERIES;
constexpr int kMaxKeys = RTE_NOALLOC_MAX_KEYS;

// Each GPU the library runs on holds its own copy until the process ends.
static __device__ __align__(256) int32_t scratch[rte::batch_scratch_words(kMaxQueries, kMaxKeys)];

// The array's device address, looked up once per device with that device
// current (the first call on it).  In host code `scratch` names the symbol
// only: kernels and copies take the looked-up address.
constexpr int kMaxDevices = 64;
std::mutex g_mu;
int32_t* g_ptr[kMaxDevices] = {};
hipError_t g_err[kMaxDevices] = {};
bool g_done[kMaxDevices] = {};

hipError_t scratch_address(int device, int32_t** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_done[device]) {
    void* ptr = nullptr;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    g_err[device] = e;
    g_ptr[device] = e == hipSuccess ? static_cast<int32_t*>(ptr) : nullptr;
    g_done[device] = true;
  }
  *out = g_ptr[device];
  return g_err[device];
}

}  // namespace

extern "C" int rte_search_batched_noalloc(rte_ctx* c, const float* PQ, int64_t nq,
                                          int64_t q_stride, const int32_t* leaf,
                                          const int32_t* k_per_q, const int32_t* beam_per_q,
                                          const rte_params* p, int k_stride, int32_t* ids_out,
                                          float* dists_out) {
  if (!c) return RTE_EINVAL;
  int32_t* base = nullptr;
  const hipError_t e = scratch_address(c->device, &base);
  if (e != hipSuccess || !base)
    return rte_fail(c, RTE_EHIP, "scratch array of device %d: %s", c->device,
                    hipGetErrorString(e));
  const rte::BatchScratch sc{base, kMaxQueries, kMaxKeys};
  return rte_search_batched_in(c, PQ, nq, q_stride, leaf, k_per_q, beam_per_q, p, k_stride,
                               ids_out, dists_out, sc);
}
