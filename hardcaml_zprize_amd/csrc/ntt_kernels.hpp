// Sub-DFT kernels shared by the single-pass and the 4-step transforms.
//
// k_dft_lds: each workgroup loads 2^log_T sequences of length L = 2^log_L
// (element i of sequence s at in[s * in_seq_stride + i * in_elem_stride],
// reduced mod p) into LDS in bit-reversed order, runs the radix-2 DIT stages of
// zprize/ntt/host/reference.cpp:109-136, optionally multiplies output k of
// sequence s by w_n^(s * k) (the 4-step twiddle, reference_model.ml:75-90) and
// by a scale, and stores element k at out[s * out_seq_stride + k * out_elem_stride].
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"
#include "gl64.hpp"
#include "ntt4096.hpp"
#include "ntt4096r16.hpp"

namespace zp {

struct DftArgs {
  const uint64_t *in;
  uint64_t *out;
  unsigned log_L, log_T;
  uint64_t in_elem_stride, in_seq_stride;
  uint64_t out_elem_stride, out_seq_stride;
  const uint64_t *tw_sub;  // w_4096^i, i < 2048
  const uint64_t *tw_full; // w_4096^i, i < 4096 (k_dft_reg's twiddle layers)
  int dir;                 // 0 forward, 1 inverse (k_dft_reg's register stages)
  const uint64_t *tw_lo;   // w_n^i, i < 4096 (4-step twiddle) or NULL
  const uint64_t *tw_hi;   // w_n^(4096 i)
  uint64_t scale;          // 1 = none
  // Batched form: when set, job blockIdx.y reads in_tab[y] and writes out_tab[y].
  const uint64_t *const *in_tab;
  uint64_t *const *out_tab;
};

static constexpr int DFT_THREADS = 256;
static constexpr unsigned DFT_LDS_LOG_ELEMS = 13;  // 64 KiB of u64

// Sequences per workgroup: at most 2^13 elements of LDS, and few enough that
// a transform still spreads over >= 2^8 workgroups (one per CU) when it has
// that many sequences; small transforms are latency-bound, and the column
// reads that a narrow tile makes uncoalesced stay in L2 at these sizes.
static inline unsigned dft_log_T(unsigned log_L, unsigned log_num_seq) {
  unsigned lt = log_L >= DFT_LDS_LOG_ELEMS ? 0 : DFT_LDS_LOG_ELEMS - log_L;
  const unsigned spread = log_num_seq > 8 ? log_num_seq - 8 : 0;
  if (lt > spread) lt = spread;
  return lt < log_num_seq ? lt : log_num_seq;
}

__device__ __forceinline__ unsigned bitrev_bits(unsigned x, unsigned bits) {
  return bits ? (__brev(x) >> (32 - bits)) : 0;
}

__global__ void __launch_bounds__(DFT_THREADS) k_dft_lds(DftArgs a) {
  extern __shared__ uint64_t tile[];
  if (a.in_tab) {
    a.in = a.in_tab[blockIdx.y];
    a.out = a.out_tab[blockIdx.y];
  }
  const unsigned L = 1u << a.log_L;
  const unsigned T = 1u << a.log_T;
  const unsigned total = L << a.log_T;
  const uint64_t seq0 = (uint64_t)xcd_tile(blockIdx.x, gridDim.x) << a.log_T;  // adjacent tiles on one XCD
  const bool in_rows = a.in_elem_stride == 1;
  for (unsigned idx = threadIdx.x; idx < total; idx += DFT_THREADS) {
    unsigned s, i;
    if (in_rows) {
      s = idx >> a.log_L;
      i = idx & (L - 1);
    } else {
      s = idx & (T - 1);
      i = idx >> a.log_T;
    }
    // Inputs >= p count as their residue (user vectors, and the lazy scratch
    // values the register-resident column kernel leaves for pass 2); the
    // stages below need canonical operands.
    const uint64_t v = gl_canon(a.in[(seq0 + s) * a.in_seq_stride + (uint64_t)i * a.in_elem_stride]);
    tile[(s << a.log_L) + bitrev_bits(i, a.log_L)] = v;
  }
  __syncthreads();
  const unsigned half = total >> 1;
  for (unsigned lm = 0; lm < a.log_L; lm++) {
    const unsigned m = 1u << lm;
    const unsigned tw_shift = 11 - lm;  // w_{2m}^k = w_4096^(k * 4096 / 2m)
    for (unsigned b = threadIdx.x; b < half; b += DFT_THREADS) {
      const unsigned s = b >> (a.log_L - 1);
      const unsigned j = b & ((L >> 1) - 1);
      const unsigned k = j & (m - 1);
      const unsigned i0 = (s << a.log_L) + ((j >> lm) << (lm + 1)) + k;
      const unsigned i1 = i0 + m;
      const uint64_t w = a.tw_sub[k << tw_shift];
      const uint64_t u = tile[i0];
      const uint64_t t = gl_mul(tile[i1], w);
      tile[i0] = gl_add(u, t);
      tile[i1] = gl_sub(u, t);
    }
    __syncthreads();
  }
  const bool out_rows = a.out_elem_stride == 1;
  for (unsigned idx = threadIdx.x; idx < total; idx += DFT_THREADS) {
    unsigned s, k;
    if (out_rows) {
      s = idx >> a.log_L;
      k = idx & (L - 1);
    } else {
      s = idx & (T - 1);
      k = idx >> a.log_T;
    }
    uint64_t v = tile[(s << a.log_L) + k];
    if (a.tw_lo) {
      const uint64_t e = (seq0 + s) * k;  // < n <= 2^24
      const uint64_t w = gl_mul(a.tw_hi[e >> 12], a.tw_lo[e & 4095]);
      v = gl_mul(v, w);
    }
    if (a.scale != 1) v = gl_mul(v, a.scale);
    a.out[(seq0 + s) * a.out_seq_stride + (uint64_t)k * a.out_elem_stride] = v;
  }
}

// ---- k_dft_reg: sub-DFTs of length L = A B C (A, B, C in {4, 8, 16}) as
// three register radix stages with LDS in between, the split of the 4096-point
// row kernel (ntt4096r16.hpp) for any L in [2^6, 2^12].  With
// j = n1 + A n2 + A B n3 and k = k3 + C k2 + C B k1,
//   X[k] = sum_n1 w_A^(n1 k1) w_L^(n1 (k3 + C k2))
//            sum_n2 w_B^(n2 k2) w_BC^(n2 k3) sum_n3 w_C^(n3 k3) x[j]:
//   stage 1  DFT_C over n3 for every (n1, n2), then x *= w_BC^(n2 k3);
//   stage 2  DFT_B over n2 for every (n1, k3), then x *= w_L^(n1 (k3 + C k2));
//   stage 3  DFT_A over n1 for every (k2, k3).
// A sequence sits in LDS as a 3-d array, element (i1, i2, i3) at
// i1 + A i2 + A B i3 (input j lands at its natural index; padded, reg_pad);
// stages 1 and 2 transform their lines in place, stage 3 writes X[k] to slot
// k.  A work item (one thread) is 16 values: 16 / R lines of a radix-R
// stage, line q L / 16 + t for item t, so that the lanes of a wave take
// consecutive lines.  The radix-R DFTs use only power-of-two twiddles (every
// root of order <= 64 is one) and the lazy arithmetic of the 4096-point
// kernels; the two twiddle layers are general products with w_4096^i from a
// table.  Inputs >= p count as their residue (the lazy stages accept any
// u64), outputs are canonical.

// Threads of a k_dft_reg workgroup: one per 16 elements of a tile of at most
// 2^13 elements.
static constexpr int DFT_REG_MAX_THREADS = 512;

template <int R>
constexpr int bitrev_r(int i) {
  int r = 0;
  for (int b = 1; b < R; b <<= 1) {
    r = (r << 1) | (i & 1);
    i >>= 1;
  }
  return r;
}
template <int R>
constexpr int log2_r() {
  return R >= 16 ? 4 : R >= 8 ? 3 : R >= 4 ? 2 : 1;
}

// x[k] = sum_m x[m] w_R^(m k) in place (R = 4, 8, 16), natural order in and
// out, any u64 in, any u64 out.
template <int R, int DIR>
__device__ __forceinline__ void dftr_regs(uint64_t (&x)[R]) {
  uint64_t y[R];
  static_for<R>([&](auto i) { y[i] = x[bitrev_r<R>(i)]; });
  static_for<R / 2>([&](auto h) { y[2 * h + 1] = gl_canon(y[2 * h + 1]); });
  static_for<log2_r<R>()>([&](auto sm1) { static_for<R / 2>([&](auto b) { dit_bfly_n<R, DIR, sm1 + 1, b>(y); }); });
  static_for<R>([&](auto i) { x[i] = y[i]; });
}

// LDS slot of element i of a sequence: one pad u64 per 16 and one more per
// 256, so that lanes 16 or 256 elements apart fall in different banks.
__device__ __forceinline__ uint32_t reg_pad(uint32_t i) { return i + (i >> 4) + (i >> 8); }

template <int LA, int LB, int LC, int DIR>
__global__ void __launch_bounds__(DFT_REG_MAX_THREADS) __attribute__((amdgpu_waves_per_eu(4))) k_dft_reg(DftArgs a) {
  constexpr int A = 1 << LA, B = 1 << LB, C = 1 << LC, LL = LA + LB + LC;
  // padded sequence stride (u64), odd so that sequences start in different banks
  constexpr uint32_t L = 1u << LL, LP = L + L / 16 + L / 256 + 1;
  extern __shared__ uint64_t tile[];
  if (a.in_tab) {
    a.in = a.in_tab[blockIdx.y];
    a.out = a.out_tab[blockIdx.y];
  }
  const unsigned T = 1u << a.log_T;
  // adjacent tiles on one XCD: a line that two tiles' 8 T-byte runs share
  // (column reads, transposed stores) meets in one L2
  const uint64_t seq0 = (uint64_t)xcd_tile(blockIdx.x, gridDim.x) << a.log_T;
  // One work item per thread: the launch gives the tile's total / 16 threads,
  // so each thread moves exactly 16 elements in and out; its 16 global loads
  // are issued together before the LDS stores.  Element offsets from the
  // tile's first element fit 32 bits (n <= 2^24).
  const bool in_rows = a.in_elem_stride == 1;
  {
    const uint64_t *in0 = a.in + seq0 * a.in_seq_stride;
    const uint32_t iss = (uint32_t)a.in_seq_stride, ies = (uint32_t)a.in_elem_stride;
    uint64_t ld[16];
    static_for<16>([&](auto r) {
      const unsigned idx = threadIdx.x + r * blockDim.x;
      const unsigned s = in_rows ? idx >> LL : idx & (T - 1), i = in_rows ? idx & (L - 1) : idx >> a.log_T;
      ld[r] = ld_u64(in0, s * iss + i * ies);
    });
    static_for<16>([&](auto r) {
      const unsigned idx = threadIdx.x + r * blockDim.x;
      const unsigned s = in_rows ? idx >> LL : idx & (T - 1), i = in_rows ? idx & (L - 1) : idx >> a.log_T;
      tile[s * LP + reg_pad(i)] = ld[r];
    });
  }
  __syncthreads();
  const unsigned w = threadIdx.x;
  const uint64_t *tw = a.tw_full;
  // stage 1: lines (n1, n2) along n3, stride A B
  {
    uint64_t *seq = tile + (w >> (LL - 4)) * LP;
    const unsigned t = w & ((1u << (LL - 4)) - 1);
    static_for<16 / C>([&](auto q) {
      const unsigned l = q * (L / 16) + t, n2 = l >> LA;
      uint64_t v[C];
      static_for<C>([&](auto e) { v[e] = seq[reg_pad(l + (e << (LA + LB)))]; });
      dftr_regs<C, DIR>(v);
      static_for<C - 1>([&](auto km1) {
        constexpr int k3 = km1 + 1;
        v[k3] = gl_mul_lz(v[k3], tw[((n2 * k3) << (12 - LB - LC)) & 4095]);
      });
      static_for<C>([&](auto e) { seq[reg_pad(l + (e << (LA + LB)))] = v[e]; });
    });
  }
  __syncthreads();
  // stage 2: lines (n1, k3) along n2, stride A
  {
    uint64_t *seq = tile + (w >> (LL - 4)) * LP;
    const unsigned t = w & ((1u << (LL - 4)) - 1);
    static_for<16 / B>([&](auto q) {
      const unsigned l = q * (L / 16) + t, n1 = l & (A - 1), k3 = l >> LA;
      const unsigned base = n1 + (k3 << (LA + LB));
      uint64_t v[B];
      static_for<B>([&](auto e) { v[e] = seq[reg_pad(base + (e << LA))]; });
      dftr_regs<B, DIR>(v);
      static_for<B>([&](auto k2) {
        v[k2] = gl_mul_lz(v[k2], tw[((n1 * (k3 + (k2 << LC))) << (12 - LL)) & 4095]);
      });
      static_for<B>([&](auto e) { seq[reg_pad(base + (e << LA))] = v[e]; });
    });
  }
  __syncthreads();
  // stage 3: lines (k2, k3) along n1, contiguous; every line is read before
  // any result is written, and X[k3 + C k2 + C B k1] goes to its natural slot
  // k (the store loop below then reads consecutive slots)
  {
    uint64_t *seq = tile + (w >> (LL - 4)) * LP;
    const unsigned t = w & ((1u << (LL - 4)) - 1);
    uint64_t v[16 / A][A];
    static_for<16 / A>([&](auto q) {
      const unsigned base = (q * (L / 16) + t) << LA;
      static_for<A>([&](auto e) { v[q][e] = seq[reg_pad(base + e)]; });
    });
    __syncthreads();
    static_for<16 / A>([&](auto q) {
      const unsigned l = q * (L / 16) + t, k2 = l & (B - 1), k3 = l >> LB;
      dftr_regs<A, DIR>(v[q]);
      static_for<A>([&](auto k1) { seq[reg_pad(k3 + (k2 << LC) + (k1 << (LB + LC)))] = v[q][k1]; });
    });
  }
  __syncthreads();
  const bool out_rows = a.out_elem_stride == 1;
  uint64_t *out0 = a.out + seq0 * a.out_seq_stride;
  const uint32_t oss = (uint32_t)a.out_seq_stride, oes = (uint32_t)a.out_elem_stride;
  // four groups of four elements (a scheduling barrier between groups keeps
  // the twiddle loads of all 16 from being hoisted at once)
  static_for<4>([&](auto g) {
    static_for<4>([&](auto rr) {
      const unsigned idx = threadIdx.x + (g * 4 + rr) * blockDim.x;
      const unsigned s = out_rows ? idx >> LL : idx & (T - 1), k = out_rows ? idx & (L - 1) : idx >> a.log_T;
      uint64_t v = tile[s * LP + reg_pad(k)];
      if (a.tw_lo) {
        const uint32_t e = (uint32_t)(seq0 + s) * k;  // < n <= 2^24
        v = gl_mul(v, gl_mul(a.tw_hi[e >> 12], a.tw_lo[e & 4095]));
      }
      v = a.scale != 1 ? gl_mul(v, a.scale) : gl_canon(v);
      st_u64(out0, s * oss + k * oes, v);
    });
    __builtin_amdgcn_sched_barrier(0);
  });
}

template <int LA, int LB, int LC>
static inline void launch_dft_reg(const DftArgs &a, dim3 grid, hipStream_t stream) {
  constexpr uint32_t L = 1u << (LA + LB + LC);
  const size_t lds = (size_t)8 * (L + L / 16 + L / 256 + 1) << a.log_T;
  const dim3 block((L << a.log_T) >> 4);  // 64 .. 512 threads: one work item each
  if (a.dir)
    hipLaunchKernelGGL((k_dft_reg<LA, LB, LC, 1>), grid, block, lds, stream, a);
  else
    hipLaunchKernelGGL((k_dft_reg<LA, LB, LC, 0>), grid, block, lds, stream, a);
}

// Sub-DFTs of 2^6 .. 2^12 points on the register-stage kernel when a
// workgroup's tile holds at least 2^10 elements (64 work items of 16 values,
// one per lane of a wave); shorter sequences and smaller tiles on the radix-2
// LDS kernel, whose 256 threads share every stage.
static inline int launch_dft(const DftArgs &a, uint64_t num_groups, hipStream_t stream, unsigned jobs = 1) {
  const dim3 grid((unsigned)num_groups, jobs);
  switch (a.log_L + a.log_T >= 10 ? a.log_L : 0) {
    case 6: launch_dft_reg<2, 2, 2>(a, grid, stream); break;
    case 7: launch_dft_reg<3, 2, 2>(a, grid, stream); break;
    case 8: launch_dft_reg<3, 3, 2>(a, grid, stream); break;
    case 9: launch_dft_reg<3, 3, 3>(a, grid, stream); break;
    case 10: launch_dft_reg<4, 3, 3>(a, grid, stream); break;
    case 11: launch_dft_reg<4, 4, 3>(a, grid, stream); break;
    case 12: launch_dft_reg<4, 4, 4>(a, grid, stream); break;
    default: {
      const size_t lds = (size_t)8 << (a.log_L + a.log_T);
      hipLaunchKernelGGL(k_dft_lds, grid, dim3(DFT_THREADS), lds, stream, a);
    }
  }
  ZP_HIP(hipGetLastError());
  return ZP_OK;
}

}  // namespace zp
