// MSM device kernels.
//
// Per msm_mult batch (N points, W windows):
//   k_digits_count   signed-digit window decomposition of every scalar
//                    (zprize/msm_pippenger/hardcaml/test/top/utils.ml:243-288,
//                    generalised to any window widths) and a histogram of
//                    (window, bucket) keys; zero digits are skipped
//                    (libs/pippenger/src/controller.ml:128-132);
//   scan             bucket offsets and work-unit offsets;
//   k_digits_scatter counting-sort placement: vals[pos] = point | sign << 31;
//   k_fill_units     unit -> bucket map (a unit is <= U entries of one bucket);
//   k_accumulate     one thread per unit: mixed twisted-Edwards additions of
//                    the unit's points into an extended accumulator
//                    (the FPGA bucket pipeline, top.ml / mixed_add_precompute.ml);
//   k_bucket_sum     per bucket: sum of its unit partials;
//   k_window_groups  per group of g buckets: running-sum reduction
//                    (driver.cpp:318-339, pippenger.h:25-33) done as
//                    sum_b (b - lo + 1) S_b + lo * sum_b S_b;
//   k_window_sum     one workgroup per window: tree sum of the group results.
// The window combine and the affine conversion run on the host (a few
// thousand field operations, the part of driver.cpp:340-375 left).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ec.hpp"

namespace zp {

static constexpr int MSM_MAXW = 32;
static constexpr int MSM_THREADS = 256;

struct MsmLayout {
  int W;
  int width[MSM_MAXW];
  int off[MSM_MAXW];
  uint32_t nb[MSM_MAXW];    // buckets of window w
  uint32_t base[MSM_MAXW];  // first global bucket index of window w
  uint32_t K;               // total buckets
};

// Bits [off, off + len) of a 253-bit scalar (bits >= 253 read as 0,
// config.ml:5-9 scalar_bits = 253).
__device__ __forceinline__ uint32_t scalar_slice(const uint64_t s[4], int off, int len) {
  uint32_t v = 0;
  const int w = off >> 6, b = off & 63;
  uint64_t lo = s[w] >> b;
  if (b && w + 1 < 4) lo |= s[w + 1] << (64 - b);
  v = (uint32_t)lo;
  int top = off + len;
  if (top > 253) len = 253 - off;
  if (len <= 0) return 0;
  return v & ((len >= 32) ? 0xffffffffu : ((1u << len) - 1));
}

// Digit of window w given the incoming carry; returns signed digit.
__device__ __forceinline__ int32_t window_digit(const uint64_t s[4], const MsmLayout &L, int w, uint32_t &carry) {
  const int width = L.width[w];
  const uint32_t slice = scalar_slice(s, L.off[w], width) + carry;
  if (w == L.W - 1) {
    carry = 0;
    return (int32_t)slice;
  }
  if (slice >= (1u << width)) {
    carry = 1;
    return 0;
  }
  if (slice >= (1u << (width - 1))) {
    carry = 1;
    return (int32_t)slice - (int32_t)(1u << width);
  }
  carry = 0;
  return (int32_t)slice;
}

__device__ __forceinline__ void load_scalar(const uint64_t *scalars, uint64_t i, uint64_t s[4]) {
  const ulonglong2 *p = reinterpret_cast<const ulonglong2 *>(scalars + 4 * i);
  ulonglong2 a = p[0], b = p[1];
  s[0] = a.x;
  s[1] = a.y;
  s[2] = b.x;
  s[3] = b.y;
}

__global__ void __launch_bounds__(MSM_THREADS) k_digits_count(const uint64_t *scalars, const uint8_t *conv, uint64_t n,
                                                              MsmLayout L, uint32_t *cnt) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    if (!conv[i]) continue;
    uint64_t s[4];
    load_scalar(scalars, i, s);
    uint32_t carry = 0;
    for (int w = 0; w < L.W; w++) {
      const int32_t d = window_digit(s, L, w, carry);
      if (d != 0) {
        const uint32_t mag = (uint32_t)(d < 0 ? -d : d);
        atomicAdd(&cnt[L.base[w] + mag - 1], 1u);
      }
    }
  }
}

__global__ void __launch_bounds__(MSM_THREADS) k_digits_scatter(const uint64_t *scalars, const uint8_t *conv,
                                                                uint64_t n, MsmLayout L, uint32_t *cursor,
                                                                uint32_t *vals) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    if (!conv[i]) continue;
    uint64_t s[4];
    load_scalar(scalars, i, s);
    uint32_t carry = 0;
    for (int w = 0; w < L.W; w++) {
      const int32_t d = window_digit(s, L, w, carry);
      if (d != 0) {
        const uint32_t mag = (uint32_t)(d < 0 ? -d : d);
        const uint32_t pos = atomicAdd(&cursor[L.base[w] + mag - 1], 1u);
        vals[pos] = (uint32_t)i | (d < 0 ? 0x80000000u : 0u);
      }
    }
  }
}

// Signed digits of every scalar for inspection (diagnostics).
__global__ void k_digits_dump(const uint64_t *scalars, uint64_t n, MsmLayout L, int32_t *out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t s[4];
    load_scalar(scalars, i, s);
    uint32_t carry = 0;
    for (int w = 0; w < L.W; w++) out[i * L.W + w] = window_digit(s, L, w, carry);
  }
}

// nunits[k] = ceil(cnt[k] / U); k == K gets 0 (so the scan's last entry is the total).
__global__ void k_units_per_bucket(const uint32_t *cnt, uint32_t K, uint32_t log_u, uint32_t *nunits) {
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k <= K; k += gridDim.x * blockDim.x) {
    nunits[k] = k < K ? (cnt[k] + (1u << log_u) - 1) >> log_u : 0;
  }
}

__global__ void k_fill_units(const uint32_t *nunits, const uint32_t *ustart, uint32_t K, uint32_t *ubucket) {
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < K; k += gridDim.x * blockDim.x) {
    const uint32_t s0 = ustart[k], nu = nunits[k];
    for (uint32_t s = 0; s < nu; s++) ubucket[s0 + s] = k;
  }
}

__device__ __forceinline__ void store_xyzt(Xyzt *dst, const Xyzt &v) { *dst = v; }

__global__ void __launch_bounds__(MSM_THREADS) k_accumulate(const PackedPre *pts, const uint32_t *vals,
                                                            const uint32_t *off, const uint32_t *cnt,
                                                            const uint32_t *ustart, const uint32_t *ubucket,
                                                            const uint32_t *total_units, uint32_t log_u,
                                                            Xyzt *partial, FConsts K) {
  const uint32_t total = *total_units;
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < total; u += gridDim.x * blockDim.x) {
    const uint32_t k = ubucket[u];
    const uint32_t s = u - ustart[k];
    const uint32_t b0 = off[k] + (s << log_u);
    const uint32_t b1 = min(off[k] + cnt[k], b0 + (1u << log_u));
    Xyzt acc = ec_identity(K);
    for (uint32_t e = b0; e < b1; e++) {
      const uint32_t v = vals[e];
      const PrePt p = pre_unpack(pts[v & 0x7fffffffu]);
      ec_madd(acc, p, (v >> 31) != 0, K);
    }
    store_xyzt(&partial[u], acc);
  }
}

__global__ void __launch_bounds__(MSM_THREADS) k_bucket_sum(const Xyzt *partial, const uint32_t *ustart,
                                                            const uint32_t *nunits, uint32_t K_total,
                                                            Xyzt *buckets, FConsts K) {
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < K_total; k += gridDim.x * blockDim.x) {
    const uint32_t nu = nunits[k], s0 = ustart[k];
    Xyzt acc = nu ? partial[s0] : ec_identity(K);
    for (uint32_t s = 1; s < nu; s++) ec_add(acc, partial[s0 + s], K);
    buckets[k] = acc;
  }
}

// acc = m * p for a small non-negative integer m (double-and-add, MSB first).
__device__ __forceinline__ Xyzt ec_mul_small(const Xyzt &p, uint32_t m, const FConsts &K) {
  Xyzt r = ec_identity(K);
  if (m == 0) return r;
  const int top = 31 - __clz(m);
  r = p;
  for (int b = top - 1; b >= 0; b--) {
    ec_dbl(r, K);
    if ((m >> b) & 1) ec_add(r, p, K);
  }
  return r;
}

// Group j of window w covers buckets [lo, lo + g) (clipped to nb[w]).
// gstart[w] = index of the first group of window w; groups[gstart[w] + j].
struct GroupLayout {
  uint32_t log_g;
  uint32_t gstart[MSM_MAXW + 1];
};

__global__ void __launch_bounds__(MSM_THREADS) k_window_groups(const Xyzt *buckets, MsmLayout L, GroupLayout G,
                                                               Xyzt *groups, FConsts K) {
  const uint32_t total = G.gstart[L.W];
  for (uint32_t gi = blockIdx.x * blockDim.x + threadIdx.x; gi < total; gi += gridDim.x * blockDim.x) {
    int w = 0;
    while (gi >= G.gstart[w + 1]) w++;
    const uint32_t j = gi - G.gstart[w];
    const uint32_t lo = j << G.log_g;
    const uint32_t hi = min(L.nb[w], lo + (1u << G.log_g));
    Xyzt run = ec_identity(K), acc = ec_identity(K);
    for (int b = (int)hi - 1; b >= (int)lo; b--) {
      ec_add(run, buckets[L.base[w] + b], K);
      ec_add(acc, run, K);
    }
    // bucket b holds digit magnitude b + 1: weight (b + 1) = (b - lo + 1) + lo
    if (lo) {
      const Xyzt t = ec_mul_small(run, lo, K);
      ec_add(acc, t, K);
    }
    groups[gi] = acc;
  }
}

// One workgroup per window: tree sum of its groups into out[w].
__global__ void __launch_bounds__(MSM_THREADS) k_window_sum(const Xyzt *groups, GroupLayout G, Xyzt *out, FConsts K) {
  extern __shared__ Xyzt sh[];
  const int w = blockIdx.x;
  const uint32_t g0 = G.gstart[w], g1 = G.gstart[w + 1];
  Xyzt acc = ec_identity(K);
  for (uint32_t g = g0 + threadIdx.x; g < g1; g += blockDim.x) ec_add(acc, groups[g], K);
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (unsigned s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      Xyzt a = sh[threadIdx.x];
      ec_add(a, sh[threadIdx.x + s], K);
      sh[threadIdx.x] = a;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[w] = sh[0];
}

}  // namespace zp
