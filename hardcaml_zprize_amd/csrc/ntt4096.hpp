// Register-resident 4096-point sub-DFTs for the 4-step transform (n1 or n2 = 4096).
//
// A workgroup holds a tile of T sequences x 4096 points entirely in VGPRs:
// thread (s, a) keeps the 64 elements j = a + 64 m (m = 0..63) of sequence s.
//   step A  radix-64 DFT over m in registers; every twiddle is a power of two
//           (w_64 = OMEGA[6] = 2^39), so it is a shift and a reduction;
//   step B  twiddle w_4096^(a k') (table [k'][a] in global memory, L1/L2
//           resident: the lanes of a wave read one contiguous run);
//   step C  transpose through LDS (32-bit halves, 2 rounds) so thread (s, k')
//           holds the 64 values over a;
//   step D  radix-64 DFT over a -> output k = k' + 64 k''.
// Pass 1 (columns of the row-major n1 x n2 matrix): loads/stores 8*T-byte row
// segments, fuses the 4-step twiddle w_n^(j2 k1) and writes the scratch
// matrix in the same layout.  Pass 2 (rows): 512-byte coalesced row loads,
// transposed store X[k2 * n1 + k1] in 8*T-byte segments.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

#include "gl64.hpp"
#include "xcd.hpp"

namespace zp {

// x * 2^S (mod p) for a compile-time S in [0, 192).
template <int S>
__device__ __forceinline__ uint64_t gl_shl(uint64_t a) {
  if constexpr (S == 0) {
    return a;
  } else if constexpr (S >= 96) {
    return gl_neg(gl_shl<S - 96>(a));
  } else if constexpr (S < 64) {
    return gl_reduce128(a << S, a >> (64 - S));
  } else if constexpr (S == 64) {
    return gl_reduce128(0, a);
  } else {
    // a 2^(S-64) fits in 128 bits; times 2^64 is that residue as a high word.
    return gl_reduce128(0, gl_reduce128(a << (S - 64), a >> (128 - S)));
  }
}

// t = v * 2^E, canonical (< p), for any u64 v and a compile-time E in [1, 96).
// Three closed forms, by where the product's 2^64 and 2^96 words fall:
//   E <= 32      v 2^E = lo + hi eps, hi = v >> (64 - E) < 2^32 (2^64 = eps);
//                after a carry the result is already < p.
//   32 < E < 64  v 2^(E-32) = lo + hi 2^64, so v 2^E = lo 2^32 - hi
//                = (lo_lo + lo_hi) 2^32 - (lo_hi + hi) = A 2^32 - B
//                = (A0 + A1 - B1) 2^32 - (A1 + B0)  (A = A1 2^32 + A0,
//                B = B1 2^32 + B0, A1 and B1 carries), in (-2^33, p): add p
//                if negative.  The 32-bit words go through one carry chain
//                (no 64-bit zero-extended temporaries, no compares).
//   E >= 64      v 2^E = -v 2^-K (K = 96 - E in (0, 32]) = M eps - q with
//                q = v >> K and M = low word of v << (32 - K); M eps < p.
template <int E>
__host__ __device__ __forceinline__ uint64_t gl_shl_lz(uint64_t v) {
  static_assert(E > 0 && E < 96, "exponent class");
  const uint32_t vl = (uint32_t)v, vh = (uint32_t)(v >> 32);
  if constexpr (E <= 32) {
    const uint64_t lo = v << E;
    const uint32_t hi = (uint32_t)(v >> (64 - E));
    const uint64_t r0 = (uint64_t)hi * GL_EPS + lo;
    const bool c = r0 < lo;
    return (c || r0 >= GL_P) ? r0 + GL_EPS : r0;
  } else if constexpr (E < 64) {
    constexpr int F = E - 32;
    const uint32_t lo_lo = vl << F;
    const uint32_t lo_hi = (uint32_t)(v >> (32 - F));
    const uint32_t hi = vh >> (32 - F);
    unsigned A1, B1, b1, b2, c2;
    const uint32_t A0 = __builtin_addc(lo_lo, lo_hi, 0u, &A1);
    const uint32_t B0 = __builtin_addc(lo_hi, hi, 0u, &B1);
    // low word -(A1 + B0) with its borrow b1; high word A0 + A1 - B1 - b1.
    // A0 + A1 never carries: A1 = 1 means lo_lo + lo_hi >= 2^32, so A0 <=
    // 2^32 - 2.  The high word's borrow b2 is then the sign of the value.
    const uint32_t r_lo = __builtin_subc(0u, B0, A1, &b1);
    const uint32_t t_hi = __builtin_addc(A0, 0u, A1, &c2);
    const uint32_t r_hi = __builtin_subc(t_hi, B1, b1, &b2);
    const uint64_t r = ((uint64_t)r_hi << 32) | r_lo;
    return b2 ? r + GL_P : r;  // + p = - eps (mod 2^64)
  } else {
    constexpr int K = 96 - E;
    const uint32_t M = K == 32 ? vl : (vl << ((32 - K) & 31));
    const uint64_t q = v >> K;
    const uint64_t me = (uint64_t)M * GL_EPS;
    const uint64_t d = me - q;
    return me < q ? d + GL_P : d;
  }
}

// Compile-time loop: f(std::integral_constant<int, I>) for I in [0, N).
template <typename Fn, int... I>
__host__ __device__ __forceinline__ void static_for_impl(Fn &&f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename Fn>
__host__ __device__ __forceinline__ void static_for(Fn &&f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

constexpr int bitrev6(int i) {
  return ((i & 1) << 5) | ((i & 2) << 3) | ((i & 4) << 1) | ((i & 8) >> 1) | ((i & 16) >> 3) | ((i & 32) >> 5);
}

constexpr int bitrev3(int i) { return ((i & 1) << 2) | (i & 2) | ((i & 4) >> 2); }

// Butterfly B (0..31) of radix-2 DIT stage S (half-size 2^(S-1)) on the
// bit-reversed array y, lazy: the twiddled operand is made canonical (by
// gl_shl_lz, or gl_canon for a unit twiddle after stage 1), which is what
// gl_add_lz / gl_sub_lz need; the other operand and the results are any u64.
template <int DIR, int S, int B>
__host__ __device__ __forceinline__ void dit_bfly(uint64_t (&y)[64]) {
  constexpr int half = 1 << (S - 1);
  constexpr int j = B % half;
  constexpr int i0 = (B / half) * 2 * half + j;
  constexpr int e0 = (39 * j * (64 >> S)) % 192;
  constexpr int e = DIR ? (192 - e0) % 192 : e0;
  // 2^e = -2^(e - 96) for e >= 96: the sign swaps the butterfly's add and sub.
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

// y[k] = sum_m x[m] w_64^(m k) in place, natural order in and out: radix-2
// DIT on the bit-reversed input, every index a compile-time constant so the
// 64 values stay in registers.  DIR 0: w_64 = 2^39; DIR 1: w_64^-1 = 2^153.
// Lazy: x[32..63] are the twiddled operands of the first stage and must be
// canonical -- CANON_IN makes them so here, class by class -- x[0..31] may be
// any u64; the results are any u64.
// Stages 1-3 only combine the eight values x[c + 8u] (u < 8) of one residue
// class c mod 8, which sit at y[8 bitrev3(c) .. + 8): they run class by class,
// c = 0..7, so when the caller has issued its loads in the same order
// (load_order) the first classes are transformed while later loads are still
// in flight.
template <int DIR, bool CANON_IN = false>
__host__ __device__ __forceinline__ void dft64_regs(uint64_t (&x)[64]) {
  uint64_t y[64];
  static_for<64>([&](auto i) { y[i] = x[bitrev6(i)]; });
  static_for<8>([&](auto c) {
    constexpr int g = bitrev3(c);
    if constexpr (CANON_IN) {
      // y[8g + t] = x[8 bitrev3(t) + c]: the upper half of the class is t odd.
      static_for<4>([&](auto h) { y[8 * g + 2 * h + 1] = gl_canon(y[8 * g + 2 * h + 1]); });
    }
    static_for<3>([&](auto sm1) {
      static_for<4>([&](auto q) { dit_bfly<DIR, sm1 + 1, 4 * g + q>(y); });
    });
  });
  static_for<3>([&](auto sm4) {
    static_for<32>([&](auto b) { dit_bfly<DIR, sm4 + 4, b>(y); });
  });
  static_for<64>([&](auto i) { x[i] = y[i]; });
}

// Calls f(m) for m = 0..63 in residue-class order (c + 8u, c outer), the
// order in which dft64_regs consumes its inputs.
template <typename Fn>
__host__ __device__ __forceinline__ void load_order(Fn &&f) {
  static_for<8>([&](auto c) { static_for<8>([&](auto u) { f(std::integral_constant<int, c + 8 * u>{}); }); });
}

// base[off] for a wave-uniform base and a per-lane element offset < 2^29:
// the address is an SGPR pair plus one 32-bit VGPR, so unrolled loops over
// a compile-time row index keep no per-row 64-bit addresses live.  The base
// may come from a job table in memory, so it is cast to the global address
// space explicitly: generic (flat) accesses would also count against the LDS
// counter and force a full wait before the first use of any load.
typedef __attribute__((address_space(1))) uint64_t gmem_u64;

__device__ __forceinline__ uint64_t ld_u64(const uint64_t *base, uint32_t off) {
  return *(const gmem_u64 *)((const char *)base + (off << 3));
}
__device__ __forceinline__ void st_u64(uint64_t *base, uint32_t off, uint64_t v) {
  *(gmem_u64 *)((char *)base + (off << 3)) = v;
}

// x[k] *= w_4096^(a k) for k = 1..63, tw = table + a.  Loaded eight at a
// time, with scheduling barriers, so the compiler does not hoist all 63 loads
// (126 VGPRs) above the multiplies.
__device__ __forceinline__ void inner_twiddle(uint64_t (&x)[64], const uint64_t *tw) {
  static_for<8>([&](auto g) {
    uint64_t w[8];
    static_for<8>([&](auto i) {
      if constexpr (g * 8 + i > 0) w[i] = tw[(g * 8 + i) * 64];
    });
    static_for<8>([&](auto i) {
      if constexpr (g * 8 + i > 0) x[g * 8 + i] = gl_mul_lz(x[g * 8 + i], w[i]);
    });
    __builtin_amdgcn_sched_barrier(0);
  });
}

// Same product with the loads one group ahead: group g + 1 is in flight while
// group g multiplies, and w0 holds group 0 (w0[0] unused), which the caller
// loads early (before the first DFT).  Used where the 16 extra VGPRs fit
// without spills (the columns kernel: -2.4 % kernel time; the rows kernel
// spills with it).
__device__ __forceinline__ void inner_twiddle_pf(uint64_t (&x)[64], const uint64_t *tw, const uint64_t (&w0)[8]) {
  uint64_t w[2][8];
  static_for<8>([&](auto i) { w[0][i] = w0[i]; });
  static_for<8>([&](auto g) {
    if constexpr (g + 1 < 8) {
      static_for<8>([&](auto i) { w[(g + 1) & 1][i] = tw[((g + 1) * 8 + i) * 64]; });
    }
    static_for<8>([&](auto i) {
      if constexpr (g * 8 + i > 0) x[g * 8 + i] = gl_mul_lz(x[g * 8 + i], w[g & 1][i]);
    });
    __builtin_amdgcn_sched_barrier(0);
  });
}

struct Ntt4096Args {
  const uint64_t *in;
  uint64_t *out;
  uint64_t n1, n2;         // matrix rows / columns (n = n1 * n2)
  const uint64_t *tw4096;  // [k][a] = w_4096^(a k), a, k < 64 (direction-specific)
  const uint64_t *tw_lo;   // w_n^i, i < 4096 (pass 1 only)
  const uint64_t *tw_hi;   // w_n^(4096 i)
  uint64_t scale;          // pass 1 of the inverse: n^-1 (unused otherwise)
  uint32_t tiles;          // number of T-wide tiles
  // Batched form: when set, job blockIdx.y reads in_tab[y] and writes out_tab[y].
  const uint64_t *const *in_tab;
  uint64_t *const *out_tab;
};


// LDS exchange index (u32 words).  The row stride 65 T and the inner
// strides are chosen so that both the writers (consecutive lanes) and the
// readers (lanes spread over (s, k')) hit distinct banks.
//   columns kernel: writers are lanes (s, a) with s fastest -> [k'][a][s];
//   rows kernel:    writers are lanes (a, s) with a fastest -> [k'][s][a].
template <int T, bool ROWS>
__device__ __forceinline__ uint32_t xidx(uint32_t kp, uint32_t a, uint32_t s) {
  return ROWS ? kp * (65 * T) + s * 65 + a : kp * (65 * T) + a * T + s;
}

template <int T>
constexpr size_t ntt4096_lds_bytes() {
  return (size_t)64 * 65 * T * sizeof(uint32_t);
}

// x[k'] of thread (s_w, a_w) -> x[a] of thread (s_r, k' = kp_r), in two
// 32-bit rounds.
template <int T, bool ROWS>
__device__ __forceinline__ void exchange(uint64_t (&x)[64], uint32_t *lds, uint32_t s_w, uint32_t a_w, uint32_t s_r,
                                         uint32_t kp_r) {
  uint32_t lo[64];
#pragma unroll
  for (int k = 0; k < 64; k++) lds[xidx<T, ROWS>(k, a_w, s_w)] = (uint32_t)x[k];
  __syncthreads();
#pragma unroll
  for (int a = 0; a < 64; a++) lo[a] = lds[xidx<T, ROWS>(kp_r, a, s_r)];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 64; k++) lds[xidx<T, ROWS>(k, a_w, s_w)] = (uint32_t)(x[k] >> 32);
  __syncthreads();
#pragma unroll
  for (int a = 0; a < 64; a++) x[a] = ((uint64_t)lds[xidx<T, ROWS>(kp_r, a, s_r)] << 32) | lo[a];
}

// A 4-wave tile (T = 4) is held to 256 VGPRs so that two workgroups share a
// CU (LDS 66.5 KB each) and one's loads and barriers overlap the other's
// arithmetic; an 8-wave tile fills the CU by itself.
template <int T>
constexpr int ntt4096_waves_per_eu() {
  return T <= 4 ? 2 : 1;
}

template <int T, int DIR>
__global__ void __launch_bounds__(64 * T) __attribute__((amdgpu_waves_per_eu(ntt4096_waves_per_eu<T>())))
k_ntt4096_cols(Ntt4096Args args) {
  extern __shared__ uint32_t lds[];
  if (args.in_tab) {
    args.in = args.in_tab[blockIdx.y];
    args.out = args.out_tab[blockIdx.y];
  }
  const uint32_t tile = xcd_tile(blockIdx.x, args.tiles);
  const uint32_t tid = threadIdx.x;
  const uint32_t s = tid % T, a = tid / T;  // column within tile, sub-sequence
  const uint64_t col = (uint64_t)tile * T + s;
  const uint64_t n2 = args.n2;
  const uint32_t lane = a * (uint32_t)n2 + s;  // element (a, s) of the tile's first row block
  uint64_t x[64];
  // Table loads that do not depend on the data go first: the 4-step twiddle
  // factors of the epilogue and the first group of inner twiddles.
  const uint64_t n = args.n1 * n2;
  const uint64_t e0 = (col * a) & (n - 1);
  const uint64_t e1 = (col * 64) & (n - 1);
  const uint64_t th0 = args.tw_hi[e0 >> 12], tl0 = args.tw_lo[e0 & 4095];
  const uint64_t th1 = args.tw_hi[e1 >> 12], tl1 = args.tw_lo[e1 & 4095];
  const uint64_t *tw = args.tw4096 + a;
  uint64_t w0[8];
  static_for<8>([&](auto i) { w0[i] = i > 0 ? tw[i * 64] : 1; });
  // All 64 loads are issued before any arithmetic (memory-level parallelism
  // for the strided row segments); the waits then come class by class.
  load_order([&](auto m) { x[m] = ld_u64(args.in + (uint64_t)(64 * m) * n2 + (uint64_t)tile * T, lane); });
  __builtin_amdgcn_sched_barrier(0);
  dft64_regs<DIR, true>(x);  // inputs >= p count as their residue
  inner_twiddle_pf(x, tw, w0);
  exchange<T, false>(x, lds, s, a, s, a);
  dft64_regs<DIR, true>(x);
  // x[k''] = output k1 = a + 64 k'' (a now plays k').  4-step twiddle
  // w_n^(col * k1) = w_n^(col * a) * (w_n^(64 col))^k''; the inverse's n^-1
  // rides on its base (one multiply per lane, none per element).  The scratch
  // matrix holds lazy (not necessarily canonical) values; pass 2
  // canonicalizes the ones its first butterfly stage needs.
  uint64_t w = gl_mul(th0, tl0);
  if constexpr (DIR == 1) w = gl_mul(w, args.scale);
  const uint64_t step = gl_mul(th1, tl1);
#pragma unroll
  for (int k = 0; k < 64; k++) {
    st_u64(args.out + (uint64_t)(64 * k) * n2 + (uint64_t)tile * T, lane, gl_mul_lz(x[k], w));
    w = gl_mul_lz(w, step);
  }
}

template <int T, int DIR>
__global__ void __launch_bounds__(64 * T) __attribute__((amdgpu_waves_per_eu(ntt4096_waves_per_eu<T>())))
k_ntt4096_rows(Ntt4096Args args) {
  extern __shared__ uint32_t lds[];
  if (args.in_tab) {
    args.in = args.in_tab[blockIdx.y];
    args.out = args.out_tab[blockIdx.y];
  }
  const uint32_t tile = xcd_tile(blockIdx.x, args.tiles);
  const uint32_t tid = threadIdx.x;
  // load mapping: a wave reads 64 consecutive elements of one row
  const uint32_t a = tid & 63, r = tid >> 6;
  const uint64_t row0 = (uint64_t)tile * T;
  const uint64_t n2 = args.n2;
  const uint32_t lane_in = r * (uint32_t)n2 + a;
  uint64_t x[64];
  load_order([&](auto m) { x[m] = ld_u64(args.in + row0 * n2 + 64 * m, lane_in); });
  __builtin_amdgcn_sched_barrier(0);
  dft64_regs<DIR, true>(x);
  inner_twiddle(x, args.tw4096 + a);
  // after the exchange thread (r2, kp) with r2 fastest, for T-row store runs
  const uint32_t r2 = tid % T, kp = tid / T;
  exchange<T, true>(x, lds, r, a, r2, kp);
  dft64_regs<DIR, true>(x);
  // Keep the stores out of the last butterfly stages: interleaving them there
  // raised the register pressure past 256 VGPRs (spills whose reloads would
  // also wait for every earlier store).
  __builtin_amdgcn_sched_barrier(0);
  const uint64_t n1 = args.n1;
  const uint32_t lane_out = kp * (uint32_t)n1 + r2;
#pragma unroll
  for (int k = 0; k < 64; k++) {
    const uint64_t v = gl_canon(x[k]);  // n^-1 of the inverse was applied in pass 1
    st_u64(args.out + (uint64_t)(64 * k) * n1 + row0, lane_out, v);
  }
}

}  // namespace zp
