// rs_peak.hip -- the device's issue rate of v_mad_u64_u32, the 32x32->64
// multiply-add the integer Montgomery products are built from (a product of
// rs_mont.h's mmul issues 8 of them, 4 v_mul_lo_u32 and 2 v_mul_hi_u32):
// the ceiling of the integer path's roofline, measured rather than assumed.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/rssieve.h"

namespace rs {
void set_error(const std::string& msg);
}

namespace {

constexpr int kChains = 8;

// Every lane runs kChains independent chains x = lo(x) * m + x, each step a
// single v_mad_u64_u32 (the 64-bit addend is the chain's own register pair);
// the XOR of the chains is stored so that no step is dead code.
__global__ __launch_bounds__(256) void k_peak_mad(uint64_t* __restrict__ out, uint32_t iters, uint32_t m) {
  const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t x[kChains];
#pragma unroll
  for (int j = 0; j < kChains; j++) x[j] = (uint64_t)gid * kChains + j + 1;
  for (uint32_t i = 0; i < iters; i++) {
#pragma unroll
    for (int j = 0; j < kChains; j++) x[j] = (uint64_t)(uint32_t)x[j] * m + x[j];
  }
  uint64_t s = 0;
#pragma unroll
  for (int j = 0; j < kChains; j++) s ^= x[j];
  out[gid] = s;
}

}  // namespace

extern "C" int rs_peak_mad_u64_u32(rs_ctx* ctx, double* lane_ops_per_s) {
  if (!ctx || !lane_ops_per_s) {
    rs::set_error("rs_peak_mad_u64_u32: null argument");
    return RS_E_ARG;
  }
  uint32_t cus = 0;
  if (rs_device_info(ctx, &cus, nullptr, nullptr, 0)) return RS_E_ARG;
  void* sv = nullptr;
  if (rs_context_stream(ctx, &sv)) return RS_E_ARG;
  hipStream_t s = (hipStream_t)sv;
  const uint32_t blocks = cus * 16, threads = 256, iters = 1u << 14;
  uint64_t* out = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = RS_OK;
  if (hipMalloc((void**)&out, (size_t)blocks * threads * 8) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    rc = RS_E_NOMEM;
  } else {
    float best_ms = 0;
    for (int rep = 0; rep < 3 && rc == RS_OK; rep++) {  // the first launch also loads the code object
      hipLaunchKernelGGL(k_peak_mad, dim3(blocks), dim3(threads), 0, s, out, iters, 0x9E3779B1u);
      if (hipEventRecord(e0, s) != hipSuccess) rc = RS_E_HIP;
      hipLaunchKernelGGL(k_peak_mad, dim3(blocks), dim3(threads), 0, s, out, iters, 0x9E3779B1u);
      if (rc == RS_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
        rc = RS_E_HIP;
      float ms = 0;
      if (rc == RS_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = RS_E_HIP;
      if (rc == RS_OK && (rep == 0 || ms < best_ms)) best_ms = ms;
    }
    if (rc == RS_OK) *lane_ops_per_s = (double)blocks * threads * iters * kChains / (best_ms * 1e-3);
    if (rc != RS_OK) rs::set_error("rs_peak_mad_u64_u32: HIP error");
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(out);
  return rc;
}
