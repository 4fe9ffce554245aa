// Issue-rate probes for the MSM roofline: lane-operations per second of
// v_mad_u64_u32 (the multiply-add the Fq multiply is built on) and, for
// comparison, of a full-rate 32-bit add.  Each thread runs 8 independent
// dependency chains so the measurement is issue-bound, not latency-bound.
#include <hip/hip_runtime.h>

#include "common.hpp"

namespace {

constexpr int PROBE_CHAINS = 8;
constexpr int PROBE_THREADS = 256;

__global__ void __launch_bounds__(PROBE_THREADS) k_probe_mad64(uint64_t *out, uint32_t b, int iters) {
  uint64_t acc[PROBE_CHAINS];
#pragma unroll
  for (int c = 0; c < PROBE_CHAINS; c++) acc[c] = threadIdx.x + c;
  for (int i = 0; i < iters; i++) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
#pragma unroll
      for (int c = 0; c < PROBE_CHAINS; c++) acc[c] = (uint64_t)(uint32_t)acc[c] * b + acc[c];
    }
  }
  uint64_t s = 0;
#pragma unroll
  for (int c = 0; c < PROBE_CHAINS; c++) s ^= acc[c];
  out[blockIdx.x * PROBE_THREADS + threadIdx.x] = s;
}

__global__ void __launch_bounds__(PROBE_THREADS) k_probe_add32(uint32_t *out, uint32_t b, int iters) {
  uint32_t acc[PROBE_CHAINS];
#pragma unroll
  for (int c = 0; c < PROBE_CHAINS; c++) acc[c] = threadIdx.x + c;
  for (int i = 0; i < iters; i++) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
#pragma unroll
      for (int c = 0; c < PROBE_CHAINS; c++) acc[c] = (acc[c] ^ b) + acc[(c + 1) % PROBE_CHAINS];
    }
  }
  uint32_t s = 0;
#pragma unroll
  for (int c = 0; c < PROBE_CHAINS; c++) s ^= acc[c];
  out[blockIdx.x * PROBE_THREADS + threadIdx.x] = s;
}

}  // namespace

// kind 0: v_mad_u64_u32 lane-ops/s (one per multiply-add);
// kind 1: full-rate 32-bit VALU lane-ops/s (one fused v_xad_u32 per step).
extern "C" int zp_probe_valu_rate(int kind, double *lane_ops_per_s) {
  ZP_CHECK(lane_ops_per_s != nullptr && (kind == 0 || kind == 1), ZP_ERR_INVALID_ARG,
           "zp_probe_valu_rate: bad argument");
  int dev = 0, cus = 0;
  ZP_HIP(hipGetDevice(&dev));
  ZP_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int blocks = cus * 8;
  const int iters = 4096;
  void *buf = nullptr;
  ZP_HIP(hipMalloc(&buf, (size_t)blocks * PROBE_THREADS * 8));
  hipEvent_t e0, e1;
  ZP_HIP(hipEventCreate(&e0));
  ZP_HIP(hipEventCreate(&e1));
  float best_ms = 1e30f;
  for (int rep = 0; rep < 4; rep++) {
    ZP_HIP(hipEventRecord(e0, 0));
    if (kind == 0)
      hipLaunchKernelGGL(k_probe_mad64, dim3(blocks), dim3(PROBE_THREADS), 0, 0, (uint64_t *)buf, 0x9e3779b9u, iters);
    else
      hipLaunchKernelGGL(k_probe_add32, dim3(blocks), dim3(PROBE_THREADS), 0, 0, (uint32_t *)buf, 0x9e3779b9u, iters);
    ZP_HIP(hipGetLastError());
    ZP_HIP(hipEventRecord(e1, 0));
    ZP_HIP(hipEventSynchronize(e1));
    float ms = 0;
    ZP_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (rep > 0 && ms < best_ms) best_ms = ms;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(buf);
  const double per_thread = (double)iters * 4 * PROBE_CHAINS;
  *lane_ops_per_s = per_thread * blocks * PROBE_THREADS / (best_ms * 1e-3);
  return ZP_OK;
}
