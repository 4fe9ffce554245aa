// 4096-point row DFTs of the 4-step transform's second pass as 16 x 16 x 16:
// 256 threads per sequence, 16 values per thread, two LDS exchanges.
//
// With k = k3 + 16 k2 + 256 k1 and j = n1 + 16 n2 + 256 n3 (all digits < 16),
// X[k] = sum_n1 w16^(n1 k1) w4096^(n1 (k3 + 16 k2))
//          sum_n2 w16^(n2 k2) w256^(n2 k3) sum_n3 w16^(n3 k3) x[j]:
//   stage 1  thread (n1, n2): DFT16 over n3, then x[k3] *= w256^(n2 k3);
//   exchange thread (n1, k3) gathers the 16 values over n2;
//   stage 2  DFT16 over n2, then x[k2] *= w4096^(n1 (k3 + 16 k2));
//   exchange thread (k2, k3) gathers the 16 values over n1;
//   stage 3  DFT16 over n1 -> X[k3 + 16 k2 + 256 k1], k1 = 0..15.
// Every w16 power is a power of two (w16 = w64^4, w64 = 2^39), so the DFT16s
// are shifts and lazy adds (dft16_regs); the two twiddle layers are general
// products with values of a w4096^i table.  Compared with the 64 x 64 split
// (64 values per thread, one exchange, k_ntt4096_rows) this costs a second
// general twiddle per element (~15 % more VALU instructions) but needs a
// quarter of the registers: at <= 64 VGPRs and 68 KiB of LDS two 16-wave
// workgroups share a CU (eight waves per SIMD instead of two).
//
// LDS layouts (32-bit words, see R16_S1/R16_S2):
//   exchange 1  [s][k3][n2][n1], sequence stride 4096 + 8;
//   exchange 2  [s][k2 16 + k3][n1] with row stride 17, sequence stride
//               256 * 17 + 8.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ntt4096.hpp"

namespace zp {

constexpr int bitrev4(int i) { return ((i & 1) << 3) | ((i & 2) << 1) | ((i & 4) >> 1) | ((i & 8) >> 3); }

// Butterfly B of radix-2 DIT stage S on a bit-reversed array of N <= 64
// values; the stage's twiddle w_(2^S)^j = w_64^(j 64 / 2^S) is the power of
// two 2^(39 j (64 >> S) mod 192), as in dit_bfly.
template <int N, int DIR, int S, int B>
__host__ __device__ __forceinline__ void dit_bfly_n(uint64_t (&y)[N]) {
  constexpr int half = 1 << (S - 1);
  constexpr int j = B % half;
  constexpr int i0 = (B / half) * 2 * half + j;
  constexpr int e0 = (39 * j * (64 >> S)) % 192;
  constexpr int e = DIR ? (192 - e0) % 192 : e0;
  constexpr int em = e % 96;
  const uint64_t u = y[i0];
  uint64_t t;
  if constexpr (em == 0) {
    t = S == 1 ? y[i0 + half] : gl_canon(y[i0 + half]);
  } else {
    t = gl_shl_lz<em>(y[i0 + half]);
  }
  if constexpr (e >= 96) {
    y[i0] = gl_sub_lz(u, t);
    y[i0 + half] = gl_add_lz(u, t);
  } else {
    y[i0] = gl_add_lz(u, t);
    y[i0 + half] = gl_sub_lz(u, t);
  }
}

// y[k] = sum_m x[m] w_16^(m k) in place, natural order in and out; lazy as
// dft64_regs: x[8..15] (the first stage's twiddled operands) must be
// canonical unless CANON_IN, the results are any u64.
template <int DIR, bool CANON_IN = false>
__host__ __device__ __forceinline__ void dft16_regs(uint64_t (&x)[16]) {
  uint64_t y[16];
  static_for<16>([&](auto i) { y[i] = x[bitrev4(i)]; });
  if constexpr (CANON_IN) static_for<8>([&](auto h) { y[2 * h + 1] = gl_canon(y[2 * h + 1]); });
  static_for<4>([&](auto sm1) { static_for<8>([&](auto b) { dit_bfly_n<16, DIR, sm1 + 1, b>(y); }); });
  static_for<16>([&](auto i) { x[i] = y[i]; });
}

struct Ntt16Args {
  const uint64_t *in;
  uint64_t *out;
  uint64_t n1, n2;          // matrix rows / columns (n = n1 * n2)
  const uint64_t *tw_full;  // w_4096^i, i < 4096 (direction-specific)
  uint32_t tiles;
  const uint64_t *const *in_tab;  // batched form: job blockIdx.y
  uint64_t *const *out_tab;
};

constexpr int R16_T = 4;  // sequences per workgroup (1024 threads)
// Exchanges as two rounds of 32-bit halves: 68 KiB of LDS per workgroup, so
// two workgroups (eight waves per SIMD at <= 64 VGPRs) share a CU and one's
// loads and barriers overlap the other's arithmetic.  Strides in 32-bit
// words, = 8 (mod 32): each half-wave of a 32-bit access then covers the
// 32 banks once.
constexpr uint32_t R16_S1 = 4096 + 8, R16_S2 = 256 * 17 + 8;
constexpr size_t r16_lds_bytes() { return (size_t)R16_T * R16_S2 * sizeof(uint32_t); }

__device__ __forceinline__ uint32_t r16_x1(uint32_t s, uint32_t k3, uint32_t n2, uint32_t n1) {
  return s * R16_S1 + k3 * 256 + n2 * 16 + n1;
}
__device__ __forceinline__ uint32_t r16_x2(uint32_t s, uint32_t k2, uint32_t k3, uint32_t n1) {
  return s * R16_S2 + (k2 * 16 + k3) * 17 + n1;
}

// x[k3] *= w256^(n2 k3) (= w4096^(16 n2 k3)), k3 = 1..15.
__device__ __forceinline__ void r16_twiddle1(uint64_t (&x)[16], const uint64_t *tw, uint32_t n2) {
  static_for<4>([&](auto g) {
    uint64_t w[4];
    static_for<4>([&](auto i) {
      if constexpr (g * 4 + i > 0) w[i] = tw[(16 * n2 * (g * 4 + i)) & 4095];
    });
    static_for<4>([&](auto i) {
      if constexpr (g * 4 + i > 0) x[g * 4 + i] = gl_mul_lz(x[g * 4 + i], w[i]);
    });
  });
}

// x[k2] *= w4096^(n1 (k3 + 16 k2)), k2 = 0..15 (none when n1 = 0).
__device__ __forceinline__ void r16_twiddle2(uint64_t (&x)[16], const uint64_t *tw, uint32_t n1, uint32_t k3) {
  static_for<4>([&](auto g) {
    uint64_t w[4];
    static_for<4>([&](auto i) { w[i] = tw[(n1 * (k3 + 16 * (g * 4 + i))) & 4095]; });
    static_for<4>([&](auto i) { x[g * 4 + i] = gl_mul_lz(x[g * 4 + i], w[i]); });
  });
}

// Both exchanges: lds[wr(i)] = x[i] for i < 16, barrier, x[i] = lds[rd(i)],
// barrier, once for the low and once for the high 32-bit halves.
template <typename WR, typename RD>
__device__ __forceinline__ void r16_exchange(uint64_t (&x)[16], uint32_t *lds, WR wr, RD rd) {
  uint32_t lo[16];
  static_for<16>([&](auto i) { lds[wr(i)] = (uint32_t)x[i]; });
  __syncthreads();
  static_for<16>([&](auto i) { lo[i] = lds[rd(i)]; });
  __syncthreads();
  static_for<16>([&](auto i) { lds[wr(i)] = (uint32_t)(x[i] >> 32); });
  __syncthreads();
  static_for<16>([&](auto i) { x[i] = (uint64_t)lds[rd(i)] << 32 | lo[i]; });
  __syncthreads();
}

// Pass 2: the 4096-point DFTs of R16_T adjacent rows, stored transposed
// (X[k2 n1 + k1]).  Loads: thread (r, j) reads row r at j + 256 m (a wave
// reads 512 contiguous bytes); after the last exchange thread (r2, jj) with
// r2 fastest, so stores are runs of R16_T elements.
template <int DIR>
__global__ void __launch_bounds__(256 * R16_T) __attribute__((amdgpu_waves_per_eu(8)))
k_ntt4096r16_rows(Ntt16Args args) {
  extern __shared__ uint32_t lds16[];
  if (args.in_tab) {
    args.in = args.in_tab[blockIdx.y];
    args.out = args.out_tab[blockIdx.y];
  }
  const uint32_t tile = xcd_tile(blockIdx.x, args.tiles);
  const uint32_t tid = threadIdx.x;
  const uint32_t j = tid & 255, r = tid >> 8;
  const uint32_t n1 = j & 15, hi = j >> 4;
  const uint64_t row0 = (uint64_t)tile * R16_T;
  const uint64_t n2c = args.n2;
  uint64_t x[16];
  const uint32_t lane_in = r * (uint32_t)n2c + j;
  static_for<16>([&](auto m) { x[m] = ld_u64(args.in + row0 * n2c + 256 * m, lane_in); });
  dft16_regs<DIR, true>(x);
  r16_twiddle1(x, args.tw_full, hi);
  r16_exchange(x, lds16, [&](auto k3) { return r16_x1(r, k3, hi, n1); },
               [&](auto n2) { return r16_x1(r, hi, n2, n1); });
  dft16_regs<DIR, true>(x);
  r16_twiddle2(x, args.tw_full, n1, hi);
  const uint32_t r2 = tid & (R16_T - 1), jj = tid / R16_T;
  r16_exchange(x, lds16, [&](auto k2) { return r16_x2(r, k2, hi, n1); },
               [&](auto m1) { return r16_x2(r2, jj >> 4, jj & 15, m1); });
  dft16_regs<DIR, true>(x);
  const uint64_t n1c = args.n1;
  const uint32_t lane_out = jj * (uint32_t)n1c + r2;
#pragma unroll
  for (int k = 0; k < 16; k++) st_u64(args.out + (uint64_t)(256 * k) * n1c + row0, lane_out, gl_canon(x[k]));
}

}  // namespace zp
