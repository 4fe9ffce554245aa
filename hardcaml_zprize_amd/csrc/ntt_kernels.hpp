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

namespace zp {

struct DftArgs {
  const uint64_t *in;
  uint64_t *out;
  unsigned log_L, log_T;
  uint64_t in_elem_stride, in_seq_stride;
  uint64_t out_elem_stride, out_seq_stride;
  const uint64_t *tw_sub;  // w_4096^i, i < 2048
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
  const uint64_t seq0 = (uint64_t)blockIdx.x << a.log_T;
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

static inline int launch_dft(const DftArgs &a, uint64_t num_groups, hipStream_t stream, unsigned jobs = 1) {
  const size_t lds = (size_t)8 << (a.log_L + a.log_T);
  hipLaunchKernelGGL(k_dft_lds, dim3((unsigned)num_groups, jobs), dim3(DFT_THREADS), lds, stream, a);
  ZP_HIP(hipGetLastError());
  return ZP_OK;
}

}  // namespace zp
