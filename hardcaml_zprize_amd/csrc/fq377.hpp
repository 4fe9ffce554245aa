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
ZP_HD uint32_t q28(int i) {
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

// ---- inversion by the binary extended Euclidean algorithm ----
// On plain 384-bit integers in 12 little-endian 32-bit words.
ZP_HD bool w12_is_one(const uint32_t (&a)[12]) {
  uint32_t acc = a[0] ^ 1u;
#pragma unroll
  for (int i = 1; i < 12; i++) acc |= a[i];
  return acc == 0;
}

ZP_HD bool w12_geq(const uint32_t (&a)[12], const uint32_t (&b)[12]) {
  bool decided = false, geq = true;
#pragma unroll
  for (int i = 11; i >= 0; i--) {
    if (!decided && a[i] != b[i]) {
      decided = true;
      geq = a[i] > b[i];
    }
  }
  return geq;
}

// a -= b (a >= b).
ZP_HD void w12_sub(uint32_t (&a)[12], const uint32_t (&b)[12]) {
  uint64_t br = 0;
#pragma unroll
  for (int i = 0; i < 12; i++) {
    const uint64_t d = (uint64_t)a[i] - b[i] - br;
    a[i] = (uint32_t)d;
    br = (d >> 32) & 1;
  }
}

// a = (a + c) / 2 with the carry out of the top word kept (c = 0 or q).
ZP_HD void w12_add_half(uint32_t (&a)[12], const uint32_t (&c)[12]) {
  uint64_t cy = 0;
#pragma unroll
  for (int i = 0; i < 12; i++) {
    const uint64_t s = (uint64_t)a[i] + c[i] + cy;
    a[i] = (uint32_t)s;
    cy = s >> 32;
  }
#pragma unroll
  for (int i = 0; i < 11; i++) a[i] = (a[i] >> 1) | (a[i + 1] << 31);
  a[11] = (a[11] >> 1) | ((uint32_t)cy << 31);
}

// a = a - b mod q for a, b < q.
ZP_HD void w12_sub_mod(uint32_t (&a)[12], const uint32_t (&b)[12], const uint32_t (&q)[12]) {
  if (!w12_geq(a, b)) {
    uint64_t cy = 0;  // a + q < 2^378: no carry out of the top word
#pragma unroll
    for (int i = 0; i < 12; i++) {
      const uint64_t s = (uint64_t)a[i] + q[i] + cy;
      a[i] = (uint32_t)s;
      cy = s >> 32;
    }
  }
  w12_sub(a, b);
}

// 1/a in the R' domain for a != 0 (mod q): the binary extended Euclidean
// algorithm on the canonical integer v = x R' gives v^-1 mod q, and one
// multiply by R'^3 turns it into x^-1 R'.  About 1,100 shift/subtract steps
// on 12-word integers instead of the ~570 multiplications of f_inv (a^(q-2));
// not constant time (the values are public).  Returns 0 for a = 0 mod q.
ZP_HD F f_inv_fast(const F &a, const FConsts &K) {
  F qf;
#pragma unroll
  for (int i = 0; i < 14; i++) qf.l[i] = q28(i);
  const Packed qp = f_pack(qf);
  const Packed ap = f_pack(f_canonical(a));
  uint32_t q[12], u[12], v[12], x1[12], x2[12], zero[12];
#pragma unroll
  for (int i = 0; i < 12; i++) {
    q[i] = qp.w[i];
    u[i] = ap.w[i];
    v[i] = qp.w[i];
    x1[i] = i == 0;
    x2[i] = 0;
    zero[i] = 0;
  }
  uint32_t nz = 0;
#pragma unroll
  for (int i = 0; i < 12; i++) nz |= u[i];
  if (!nz) return f_zero();
  // Invariants: x1 a = u, x2 a = v (mod q); gcd(u, v) = 1 since q is prime.
  // Every step halves u or v or replaces the larger by the difference, so
  // u + v < 2^379 falls to 2 within 2 * 379 halvings; the bounds below are
  // never reached for a != 0.
  for (int it = 0; it < 1600 && !w12_is_one(u) && !w12_is_one(v); it++) {
    for (int k = 0; k < 384 && !(u[0] & 1); k++) w12_add_half(u, zero), w12_add_half(x1, (x1[0] & 1) ? q : zero);
    for (int k = 0; k < 384 && !(v[0] & 1); k++) w12_add_half(v, zero), w12_add_half(x2, (x2[0] & 1) ? q : zero);
    if (w12_is_one(u) || w12_is_one(v)) break;
    if (w12_geq(u, v)) {
      w12_sub(u, v);
      w12_sub_mod(x1, x2, q);
    } else {
      w12_sub(v, u);
      w12_sub_mod(x2, x1, q);
    }
  }
  Packed rp;
  const bool u1 = w12_is_one(u);
#pragma unroll
  for (int i = 0; i < 12; i++) rp.w[i] = u1 ? x1[i] : x2[i];
  return f_mul(f_unpack(rp), K.r3);
}

}  // namespace zp
