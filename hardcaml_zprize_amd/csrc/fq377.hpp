// BLS12-377 base field Fq on gfx950.
//
// Working form "F": 14 limbs of 28 bits in 32-bit registers, Montgomery form
// with R' = 2^392 (= 2^(28*14)).  A 28-bit radix lets the Montgomery multiply
// accumulate every 32x32->64 partial product of a column in one 64-bit
// register (at most 28 products < 2^58 each) with no carry chains: each
// product is one v_mad_u64_u32 whose 64-bit addend is the column
// accumulator, and carries move once per row.  q = 1 (mod 2^28), so the row
// reduction factor is m = -t0 (mod 2^28), no multiply.
//
// Bounds ("value < k q" is tracked at call sites):
//   f_mul: inputs with values < 2^384 and limbs < 2^29 (or one input < 2^30
//          and the other normalized, < 2^28) give a normalized output (limbs
//          < 2^28) with value < 2q: each column sums <= 28 products < 2^58,
//          and the CIOS result is (a b + m q) / R' < a b / R' + q
//          <= 2^768 / 2^392 + q < 2q.
//   fadd: normalized inputs, normalized output, value = a + b.
//   fsub_k: a + k q - b, normalized; needs b < k q (k in {2, 4, 8}).
//
// Storage form "packed": the canonical value (< q) of the R' Montgomery form
// in 12 little-endian 32-bit words (48 bytes).  The wire form of the
// reference (arkworks, R = 2^384, rust_types.h:63-68,
// test_fpga_harness/README.md:13-17) converts with one multiply by a
// constant each way (to_internal / to_wire).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zp {

#define ZP_HD __host__ __device__ __forceinline__

static constexpr uint32_t F_MASK = (1u << 28) - 1;

struct F {
  uint32_t l[14];
};

struct Packed {
  uint32_t w[12];
};

// q in radix 2^28.
__host__ __device__ __forceinline__ constexpr uint32_t q28(int i) {
  switch (i) {
    case 0: return 0x1u;
    case 1: return 0x8c0000u;
    case 2: return 0x85u;
    case 3: return 0x5d44300u;
    case 4: return 0x800170bu;
    case 5: return 0x2fba094u;
    case 6: return 0xf1ef362u;
    case 7: return 0xf5138u;
    case 8: return 0xa22d9f3u;
    case 9: return 0xa1493b1u;
    case 10: return 0xb05c06cu;
    case 11: return 0x10eac63u;
    case 12: return 0xa4617c5u;
    default: return 0x1ae3u;
  }
}

ZP_HD F f_zero() {
  F r;
#pragma unroll
  for (int i = 0; i < 14; i++) r.l[i] = 0;
  return r;
}

// Normalize limbs (< 2^32 each, non-negative) so every limb is < 2^28.
ZP_HD void f_carry(F &a) {
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < 13; i++) {
    uint32_t v = a.l[i] + c;
    a.l[i] = v & F_MASK;
    c = v >> 28;
  }
  a.l[13] += c;
}

// Limbwise a + b without carries: normalized inputs give limbs < 2^29.
ZP_HD F f_add_nc(const F &a, const F &b) {
  F r;
#pragma unroll
  for (int i = 0; i < 14; i++) r.l[i] = a.l[i] + b.l[i];
  return r;
}

// Limbwise a + K q - b without carries (value a + K q - b).  Needs a with
// limbs < 2^29, b normalized with b < K q / 2 so that every limb of the
// redundant K q (see f_make_kq) covers b's limb, top limb included; gives
// limbs < 2^30.
ZP_HD F f_sub_nc(const F &a, const F &b, const uint32_t (&kq)[14]) {
  F r;
#pragma unroll
  for (int i = 0; i < 14; i++) r.l[i] = a.l[i] + kq[i] - b.l[i];
  return r;
}

ZP_HD F f_add(const F &a, const F &b) {
  F r;
#pragma unroll
  for (int i = 0; i < 14; i++) r.l[i] = a.l[i] + b.l[i];
  f_carry(r);
  return r;
}

// a + K q - b for normalized a, b with b < K q; K q given as 14 limbs each
// >= 2^28 - 1 except the top (borrow-free per limb), see f_kq.
ZP_HD F f_sub_kq(const F &a, const F &b, const uint32_t (&kq)[14]) {
  F r;
  int32_t c = 0;
#pragma unroll
  for (int i = 0; i < 13; i++) {
    int32_t v = (int32_t)(a.l[i] + kq[i] - b.l[i]) + c;
    r.l[i] = (uint32_t)v & F_MASK;
    c = v >> 28;  // arithmetic: -1, 0 or small positive
  }
  r.l[13] = (uint32_t)((int32_t)(a.l[13] + kq[13] - b.l[13]) + c);
  return r;
}

// K*q in a redundant radix-2^28 form: limbs 0..12 lie in [2^28, 2^29) so that
// kq[i] - b[i] >= 0 for any normalized b, and the top limb absorbs the rest.
struct KQ {
  uint32_t l[14];
};

ZP_HD void f_make_kq(uint32_t k, uint32_t (&out)[14]) {
  // Plain normalized limbs of k*q (k <= 16 keeps everything in 32 bits).
  uint32_t v[14];
  uint32_t c = 0;
  for (int i = 0; i < 14; i++) {
    uint32_t t = q28(i) * k + c;
    v[i] = t & F_MASK;
    c = t >> 28;
  }
  // Borrow 1 from each next limb: limb i gets + 2^28, limb i+1 gets - 1.
  out[0] = v[0] + (1u << 28);
  for (int i = 1; i < 13; i++) out[i] = v[i] + (1u << 28) - 1;
  out[13] = v[13] - 1 + (c << 28);
}

// Limb i of K q in the borrow-free form of f_make_kq, as a compile-time
// constant (the subtractions below then take immediates instead of a table).
__host__ __device__ constexpr uint32_t kq_limb(uint32_t k, int i) {
  uint32_t v[14] = {};
  uint32_t c = 0;
  for (int j = 0; j < 14; j++) {
    const uint32_t t = q28(j) * k + c;
    v[j] = t & F_MASK;
    c = t >> 28;
  }
  return i == 0 ? v[0] + (1u << 28) : i < 13 ? v[i] + (1u << 28) - 1 : v[13] - 1 + (c << 28);
}

// f_sub_kq / f_sub_nc with K q a compile-time constant.
template <uint32_t KQ>
ZP_HD F f_sub_kq_c(const F &a, const F &b) {
  F r;
  int32_t c = 0;
#pragma unroll
  for (int i = 0; i < 13; i++) {
    int32_t v = (int32_t)(a.l[i] + kq_limb(KQ, i) - b.l[i]) + c;
    r.l[i] = (uint32_t)v & F_MASK;
    c = v >> 28;
  }
  r.l[13] = (uint32_t)((int32_t)(a.l[13] + kq_limb(KQ, 13) - b.l[13]) + c);
  return r;
}

template <uint32_t KQ>
ZP_HD F f_sub_nc_c(const F &a, const F &b) {
  F r;
#pragma unroll
  for (int i = 0; i < 14; i++) r.l[i] = a.l[i] + kq_limb(KQ, i) - b.l[i];
  return r;
}

// Montgomery product a * b / 2^392 mod q (CIOS, lazy carries).
ZP_HD F f_mul(const F &a, const F &b) {
  uint64_t t[14];
#pragma unroll
  for (int j = 0; j < 14; j++) t[j] = 0;
#pragma unroll
  for (int i = 0; i < 14; i++) {
    const uint32_t bi = b.l[i];
#pragma unroll
    for (int j = 0; j < 14; j++) t[j] += (uint64_t)a.l[j] * bi;
    const uint32_t m = (0u - (uint32_t)t[0]) & F_MASK;
    t[0] += m;  // q0 = 1
#pragma unroll
    for (int j = 1; j < 14; j++) t[j] += (uint64_t)m * q28(j);
    // t[0] is now divisible by 2^28: shift the row down by one limb.
    const uint64_t c = t[0] >> 28;
#pragma unroll
    for (int j = 0; j < 13; j++) t[j] = t[j + 1];
    t[0] += c;
    t[13] = 0;
  }
  F r;
  uint64_t c = 0;
#pragma unroll
  for (int j = 0; j < 14; j++) {
    const uint64_t v = t[j] + c;
    r.l[j] = (uint32_t)v & F_MASK;
    c = v >> 28;
  }
  return r;
}

ZP_HD F f_sqr(const F &a) { return f_mul(a, a); }

// Subtract q while the value is >= q (inputs < 16 q), giving the canonical value.
ZP_HD bool f_geq_q(const F &a) {
  // Compare normalized a with q from the top limb down.
  bool decided = false, geq = true;
#pragma unroll
  for (int i = 13; i >= 0; i--) {
    if (!decided && a.l[i] != q28(i)) {
      decided = true;
      geq = a.l[i] > q28(i);
    }
  }
  return geq;
}

ZP_HD void f_sub_q(F &a) {
  int32_t c = 0;
#pragma unroll
  for (int i = 0; i < 13; i++) {
    int32_t v = (int32_t)(a.l[i] - q28(i)) + c;
    a.l[i] = (uint32_t)v & F_MASK;
    c = v >> 28;
  }
  a.l[13] = (uint32_t)((int32_t)(a.l[13] - q28(13)) + c);
}

ZP_HD F f_canonical(F a) {
  for (int k = 0; k < 16 && f_geq_q(a); k++) f_sub_q(a);
  return a;
}

ZP_HD bool f_is_zero_canon(const F &a) {
  F c = f_canonical(a);
  uint32_t acc = 0;
#pragma unroll
  for (int i = 0; i < 14; i++) acc |= c.l[i];
  return acc == 0;
}

ZP_HD bool f_eq(const F &a, const F &b) {
  F x = f_canonical(a), y = f_canonical(b);
  uint32_t acc = 0;
#pragma unroll
  for (int i = 0; i < 14; i++) acc |= x.l[i] ^ y.l[i];
  return acc == 0;
}

// ---- packing (12 x 32-bit words <-> 14 x 28-bit limbs) ----
ZP_HD F f_unpack(const Packed &p) {
  F r;
#pragma unroll
  for (int i = 0; i < 14; i++) {
    const int bit = 28 * i;
    const int w = bit >> 5, s = bit & 31;
    uint64_t lo = p.w[w];
    uint64_t hi = (w + 1 < 12) ? p.w[w + 1] : 0;
    r.l[i] = (uint32_t)(((hi << 32) | lo) >> s) & F_MASK;
  }
  return r;
}

// a must be normalized and < 2^384 (canonical for storage).
ZP_HD Packed f_pack(const F &a) {
  Packed p;
#pragma unroll
  for (int i = 0; i < 12; i++) p.w[i] = 0;
#pragma unroll
  for (int i = 0; i < 14; i++) {
    const int bit = 28 * i;
    const int w = bit >> 5, s = bit & 31;
    const uint64_t v = (uint64_t)a.l[i] << s;
    p.w[w] |= (uint32_t)v;
    if (w + 1 < 12) p.w[w + 1] |= (uint32_t)(v >> 32);
  }
  return p;
}

// ---- constants (values in R' Montgomery form unless noted) ----
// Built on the host by the library from plain integers; see msm.hip.
struct FConsts {
  F one;        // R' mod q
  F two;        // 2 R'
  F to_int;     // 2^400 mod q: f_mul(x R, to_int) = x R'
  F r2;         // R'^2 = 2^784 mod q: f_mul(x, r2) = x R' for a plain (canonical) x
  F to_wire;    // 2^384 mod q as a plain value: f_mul(x R', to_wire) = x R
  F r3;         // R'^3 mod q as a plain value: f_mul(v^-1, r3) = x^-1 R' for v = x R' (f_inv_fast)
  F k2d;        // 2d (twisted-Edwards k of hwcd-3)
  F d2;         // 2d again, used for the precomputed t coordinate
  F mont_s;     // Weierstrass -> Montgomery scale s
  F ts;         // twisted_scale
  F alpha;      // Weierstrass alpha (= -1)
  F inv_cb;     // 1 / c_B
  F ca_3cb;     // c_A / (3 c_B)
  uint32_t q2[14], q4[14], q8[14];  // K q in borrow-free form
};

// a^(q-2) by square-and-multiply over the exponent bits of q - 2.
ZP_HD F f_inv(const F &a, const F &one) {
  // q - 2 in 32-bit words (q0 = 1 so limb 0 wraps and limb 1 borrows).
  const uint32_t e[12] = {0xffffffffu, 0x8508bfffu, 0x30000000u, 0x170b5d44u, 0xba094800u, 0x1ef3622fu,
                          0x00f5138fu, 0x1a22d9f3u, 0x6ca1493bu, 0xc63b05c0u, 0x17c510eau, 0x01ae3a46u};
  F r = one;
  for (int i = 11; i >= 0; i--) {
    for (int b = 31; b >= 0; b--) {
      r = f_sqr(r);
      if ((e[i] >> b) & 1) r = f_mul(r, a);
    }
  }
  return r;
}

// ---- inversion: binary GCD with batched divsteps ----
// T. Pornin, "Optimized Binary GCD for Modular Inversion" (IACR ePrint
// 2020/972), Algorithm 2 with k = 32: each outer step runs 31 binary-GCD
// steps on 64-bit approximations of a and b (low 31 bits and top 33 bits),
// collecting the update factors f0, g0, f1, g1 (|.| <= 2^31), then applies
// them to the full values: a' = (a f0 + b g0) / 2^31, b' = (a f1 + b g1) / 2^31
// (exact), and likewise u', v' with the division by 2^31 done mod q by adding
// k q (q = 1 mod 2^46, so k = -t mod 2^31).  Invariants a = u y, b = v y
// (mod q); 25 outer steps (775 >= 2 * 377 - 1 divsteps) end with b = 1, so
// v = y^-1.  Values are 13 limbs of 30 bits (signed 64-bit accumulators).
static constexpr int INV_L = 13;
static constexpr uint32_t INV_M30 = (1u << 30) - 1;

// r = (x f + y g) as a signed 13-limb value plus a signed top carry.
ZP_HD int64_t inv_lin(const uint32_t (&x)[INV_L], const uint32_t (&y)[INV_L], int64_t f, int64_t g,
                      uint32_t (&r)[INV_L]) {
  int64_t c = 0;
#pragma unroll
  for (int i = 0; i < INV_L; i++) {
    c += (int64_t)x[i] * f + (int64_t)y[i] * g;
    r[i] = (uint32_t)c & INV_M30;
    c >>= 30;  // arithmetic
  }
  return c;
}

// r = -r (value r + top 2^390) when top < 0; returns whether it negated.
ZP_HD bool inv_abs(uint32_t (&r)[INV_L], int64_t &top) {
  if (top >= 0) return false;
  int64_t c = 0;
#pragma unroll
  for (int i = 0; i < INV_L; i++) {
    c -= (int64_t)r[i];
    r[i] = (uint32_t)c & INV_M30;
    c >>= 30;
  }
  top = -top + c;
  return true;
}

// r = (r + top 2^390) >> 31 for a non-negative value (r[0] bits 0..29 and
// r[1] bit 0 are zero).
ZP_HD void inv_shr31(uint32_t (&r)[INV_L], int64_t top) {
#pragma unroll
  for (int i = 0; i < INV_L - 2; i++) r[i] = (r[i + 1] >> 1) | ((r[i + 2] & 1u) << 29);
  r[INV_L - 2] = (r[INV_L - 1] >> 1) | (((uint32_t)top & 1u) << 29);
  r[INV_L - 1] = (uint32_t)(top >> 1);
}

ZP_HD bool inv_geq(const uint32_t (&a)[INV_L], const uint32_t (&b)[INV_L]) {
  bool decided = false, geq = true;
#pragma unroll
  for (int i = INV_L - 1; i >= 0; i--) {
    if (!decided && a[i] != b[i]) {
      decided = true;
      geq = a[i] > b[i];
    }
  }
  return geq;
}

ZP_HD void inv_addsub(uint32_t (&a)[INV_L], const uint32_t (&b)[INV_L], bool sub) {
  int64_t c = 0;
#pragma unroll
  for (int i = 0; i < INV_L; i++) {
    c += sub ? (int64_t)a[i] - (int64_t)b[i] : (int64_t)a[i] + (int64_t)b[i];
    a[i] = (uint32_t)c & INV_M30;
    c >>= 30;
  }
}

// (x f + y g) / 2^31 mod q, for x, y in [0, q).
ZP_HD void inv_lin_mod(const uint32_t (&x)[INV_L], const uint32_t (&y)[INV_L], int64_t f, int64_t g,
                       const uint32_t (&q)[INV_L], uint32_t (&r)[INV_L]) {
  int64_t top = inv_lin(x, y, f, g, r);
  // + k q with k = -(r mod 2^31) mod 2^31 makes the low 31 bits zero
  const int64_t k = (int64_t)((0u - (r[0] | ((r[1] & 1u) << 30))) & 0x7fffffffu);
  int64_t c = 0;
#pragma unroll
  for (int i = 0; i < INV_L; i++) {
    c += (int64_t)r[i] + (int64_t)q[i] * k;
    r[i] = (uint32_t)c & INV_M30;
    c >>= 30;
  }
  top += c;
  // the value (|.| < 3 q after the shift) as sign and magnitude
  const bool neg = inv_abs(r, top);
  inv_shr31(r, top);
  if (neg) {  // r = q - |r| mod q, |r| < 3q
    for (int t = 0; t < 3 && inv_geq(r, q); t++) inv_addsub(r, q, true);
    bool z = true;
#pragma unroll
    for (int i = 0; i < INV_L; i++) z = z && r[i] == 0;
    if (!z) {
      uint32_t m[INV_L];
#pragma unroll
      for (int i = 0; i < INV_L; i++) m[i] = q[i];
      inv_addsub(m, r, true);
#pragma unroll
      for (int i = 0; i < INV_L; i++) r[i] = m[i];
    }
  } else {
    for (int t = 0; t < 3 && inv_geq(r, q); t++) inv_addsub(r, q, true);
  }
}

ZP_HD int inv_bitlen(const uint32_t (&a)[INV_L]) {
  int len = 0;
#pragma unroll
  for (int i = 0; i < INV_L; i++)
    if (a[i]) len = 30 * i + 32 - __builtin_clz(a[i]);
  return len;
}

// Bits [pos, pos + 33) of a (pos >= 0), as an integer < 2^33.
ZP_HD uint64_t inv_bits33(const uint32_t (&a)[INV_L], int pos) {
  const int l = pos / 30, s = pos % 30;
  uint64_t w = 0;
#pragma unroll
  for (int j = 0; j < 3; j++) {
    const int i = l + j;
    const uint64_t v = (i < INV_L) ? a[i] : 0;
    w |= v << (30 * j);
  }
  return (w >> s) & ((1ull << 33) - 1);
}

// 1/a in the R' domain for a != 0 (mod q): with y = x R' the canonical
// integer, y^-1 R'^3 / R' = x^-1 R'.  Sets out = 0 for a = 0 mod q.  Returns
// false if the 25 outer steps did not reach b = 1 (not expected for any a;
// the caller then falls back to f_inv).
ZP_HD bool f_inv_pornin(const F &a, const FConsts &K, F &out) {
  uint32_t q[INV_L], A[INV_L], B[INV_L], U[INV_L], V[INV_L];
  {
    F qf;
#pragma unroll
    for (int i = 0; i < 14; i++) qf.l[i] = q28(i);
    const F ac = f_canonical(a);
    // 28-bit limbs -> 30-bit limbs
#pragma unroll
    for (int i = 0; i < INV_L; i++) {
      const int bit = 30 * i, l = bit / 28, s = bit % 28;
      const uint64_t w0 = (uint64_t)ac.l[l] | ((l + 1 < 14) ? (uint64_t)ac.l[l + 1] << 28 : 0) |
                          ((l + 2 < 14) ? (uint64_t)ac.l[l + 2] << 56 : 0);
      const uint64_t w1 = (uint64_t)qf.l[l] | ((l + 1 < 14) ? (uint64_t)qf.l[l + 1] << 28 : 0) |
                          ((l + 2 < 14) ? (uint64_t)qf.l[l + 2] << 56 : 0);
      A[i] = (uint32_t)(w0 >> s) & INV_M30;
      q[i] = (uint32_t)(w1 >> s) & INV_M30;
      B[i] = q[i];
      U[i] = i == 0;
      V[i] = 0;
    }
  }
  bool nz = false;
#pragma unroll
  for (int i = 0; i < INV_L; i++) nz = nz || A[i] != 0;
  if (!nz) {
    out = f_zero();
    return true;
  }
  for (int it = 0; it < 25; it++) {
    const int la = inv_bitlen(A), lb = inv_bitlen(B);
    const int n = la > lb ? (la > 64 ? la : 64) : (lb > 64 ? lb : 64);
    uint64_t a_ = (uint64_t)(A[0] | ((A[1] & 1u) << 30)) | (inv_bits33(A, n - 33) << 31);
    uint64_t b_ = (uint64_t)(B[0] | ((B[1] & 1u) << 30)) | (inv_bits33(B, n - 33) << 31);
    int64_t f0 = 1, g0 = 0, f1 = 0, g1 = 1;
    for (int j = 0; j < 31; j++) {
      const uint64_t odd = 0ull - (a_ & 1);
      const uint64_t sw = odd & (0ull - (uint64_t)(a_ < b_));
      // conditional swap (a_, b_), (f0, f1), (g0, g1)
      const uint64_t tab = (a_ ^ b_) & sw;
      a_ ^= tab;
      b_ ^= tab;
      const int64_t tf = (f0 ^ f1) & (int64_t)sw, tg = (g0 ^ g1) & (int64_t)sw;
      f0 ^= tf;
      f1 ^= tf;
      g0 ^= tg;
      g1 ^= tg;
      // conditional subtract
      a_ -= b_ & odd;
      f0 -= f1 & (int64_t)odd;
      g0 -= g1 & (int64_t)odd;
      a_ >>= 1;
      f1 *= 2;  // (a left shift of a negative value is undefined before C++20)
      g1 *= 2;
    }
    uint32_t nA[INV_L], nB[INV_L];
    int64_t ta = inv_lin(A, B, f0, g0, nA), tb = inv_lin(A, B, f1, g1, nB);
    if (inv_abs(nA, ta)) f0 = -f0, g0 = -g0;
    if (inv_abs(nB, tb)) f1 = -f1, g1 = -g1;
    inv_shr31(nA, ta);
    inv_shr31(nB, tb);
    uint32_t nU[INV_L], nV[INV_L];
    inv_lin_mod(U, V, f0, g0, q, nU);
    inv_lin_mod(U, V, f1, g1, q, nV);
#pragma unroll
    for (int i = 0; i < INV_L; i++) {
      A[i] = nA[i];
      B[i] = nB[i];
      U[i] = nU[i];
      V[i] = nV[i];
    }
  }
  bool one = B[0] == 1;
#pragma unroll
  for (int i = 1; i < INV_L; i++) one = one && B[i] == 0;
  if (!one) return false;
  // 30-bit limbs -> 28-bit limbs of v = y^-1, then * R'^3 / R'
  F v;
#pragma unroll
  for (int i = 0; i < 14; i++) {
    const int bit = 28 * i, l = bit / 30, s = bit % 30;
    const uint64_t w = (uint64_t)V[l] | ((l + 1 < INV_L) ? (uint64_t)V[l + 1] << 30 : 0);
    v.l[i] = (uint32_t)(w >> s) & F_MASK;
  }
  out = f_mul(v, K.r3);
  return true;
}

ZP_HD F f_inv_fast(const F &a, const FConsts &K) {
  F r;
  if (!f_inv_pornin(a, K, r)) r = f_inv(a, K.one);
  return r;
}

}  // namespace zp
