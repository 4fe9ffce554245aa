// Goldilocks field p = 2^64 - 2^32 + 1 on gfx950.
//
// Canonical representatives in [0, p) at every function boundary, so results
// are bit-identical to the reference's GF class (zprize/ntt/host/gf.h:19-58).
// The multiply reduces a 128-bit product with 2^64 = 2^32 - 1 and
// 2^96 = -1 (mod p); multiplication by a power of two (the roots of unity of
// order <= 64 are all powers of two: OMEGA[6] = 2^39) uses shifts only.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zp {

static constexpr uint64_t GL_P = 0xFFFFFFFF00000001ull;
static constexpr uint64_t GL_EPS = 0xFFFFFFFFull;

__host__ __device__ __forceinline__ uint64_t gl_add(uint64_t a, uint64_t b) {
  uint64_t s = a + b;
  // On overflow the true sum is s + 2^64; s - p (mod 2^64) is then s + eps,
  // which is the reduced value.  Otherwise subtract p iff s >= p.
  bool ovf = s < a;
  return (ovf || s >= GL_P) ? s - GL_P : s;
}

__host__ __device__ __forceinline__ uint64_t gl_sub(uint64_t a, uint64_t b) {
  uint64_t d = a - b;
  return (a < b) ? d + GL_P : d;
}

__host__ __device__ __forceinline__ uint64_t gl_neg(uint64_t a) { return a ? GL_P - a : 0; }

// Reduce hi * 2^64 + lo (any 128-bit value) to [0, p).
__host__ __device__ __forceinline__ uint64_t gl_reduce128(uint64_t lo, uint64_t hi) {
  uint64_t hi_hi = hi >> 32;
  uint64_t hi_lo = hi & GL_EPS;
  // lo - hi_hi (hi_hi * 2^96 = -hi_hi); a borrow adds 2^64 = eps back as -eps.
  uint64_t t0 = lo - hi_hi;
  if (lo < hi_hi) t0 -= GL_EPS;
  // hi_lo * 2^64 = hi_lo * eps = (hi_lo << 32) - hi_lo, < 2^64.
  uint64_t t1 = (hi_lo << 32) - hi_lo;
  uint64_t r = t0 + t1;
  if (r < t1) r += GL_EPS;
  return r >= GL_P ? r - GL_P : r;
}

// 64 x 64 -> 128-bit product from four 32 x 32 -> 64 multiply-adds (each
// partial product plus one 32-bit carry word fits in 64 bits).
__host__ __device__ __forceinline__ void mul_64x64(uint64_t a, uint64_t b, uint64_t &lo, uint64_t &hi) {
  const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
  const uint32_t b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
  const uint64_t p0 = (uint64_t)a0 * b0;
  const uint64_t p1 = (uint64_t)a0 * b1 + (p0 >> 32);
  const uint64_t p2 = (uint64_t)a1 * b0 + (uint32_t)p1;
  const uint64_t p3 = (uint64_t)a1 * b1 + (p1 >> 32) + (p2 >> 32);
  lo = (p2 << 32) | (uint32_t)p0;
  hi = p3;
}

__device__ __forceinline__ uint64_t gl_mul(uint64_t a, uint64_t b) {
  uint64_t lo, hi;
  mul_64x64(a, b, lo, hi);
  return gl_reduce128(lo, hi);
}

// a * 2^s for 0 <= s < 192 (2^96 = -1).
__device__ __forceinline__ uint64_t gl_mul_pow2(uint64_t a, unsigned s) {
  bool negate = false;
  if (s >= 96) {
    s -= 96;
    negate = true;
  }
  uint64_t r;
  if (s == 0) {
    r = a >= GL_P ? a - GL_P : a;
  } else if (s < 64) {
    r = gl_reduce128(a << s, a >> (64 - s));
  } else {
    // a * 2^(s-64) fits in 128 bits (s - 64 < 32); then times 2^64 is the
    // 128-bit value with that residue as its high word.
    unsigned t = s - 64;
    uint64_t b = t == 0 ? (a >= GL_P ? a - GL_P : a) : gl_reduce128(a << t, a >> (64 - t));
    r = gl_reduce128(0, b);
  }
  return negate ? gl_neg(r) : r;
}

// ---- lazy forms: operands and results are any u64 standing for its residue
// mod p; only the arguments named "t" below must be canonical.  They let the
// 4096-point kernels skip the final compare-and-subtract on most values and
// canonicalize once, where a bound needs it or before a value is stored.

__host__ __device__ __forceinline__ uint64_t gl_canon(uint64_t a) { return a >= GL_P ? a - GL_P : a; }

// u + t, t < p.  u + t < 2^64 + p, so after a carry s < p - 1 and s + eps
// cannot carry again.
__host__ __device__ __forceinline__ uint64_t gl_add_lz(uint64_t u, uint64_t t) {
  const uint64_t s = u + t;
  const uint32_t c = s < t;
  return s + (uint64_t)c * GL_EPS;
}

// u - t, t < p.  After a borrow d = u - t + 2^64 >= 2^64 - p + 1 = eps, so
// d - eps (the 2^64 = eps fold) cannot borrow again.
__host__ __device__ __forceinline__ uint64_t gl_sub_lz(uint64_t u, uint64_t t) {
  const uint64_t d = u - t;
  const uint32_t b = u < t;
  return d - (uint64_t)b * GL_EPS;
}

// hi * 2^64 + lo reduced to a u64 of the same residue (not canonical):
// lo + hl * eps - hh with hi = hh 2^32 + hl (2^64 = eps, 2^96 = -1).
__host__ __device__ __forceinline__ uint64_t gl_reduce128_lz(uint64_t lo, uint64_t hi) {
  const uint32_t hl = (uint32_t)hi, hh = (uint32_t)(hi >> 32);
  const uint64_t r0 = (uint64_t)hl * GL_EPS + lo;
  const bool c = r0 < lo;  // r0 stands for r0 + 2^64 = r0 + eps
  const uint64_t r1 = r0 - hh;
  const bool b = r0 < hh;  // r1 stands for r1 - 2^64 = r1 - eps
  // c only: r0 < (2^32 - 1)^2, so r1 + eps < 2^64.  b only: r1 >= 2^64 - 2^32,
  // so r1 - eps >= 0.  Both or neither: r1 as it is.
  uint64_t r = r1;
  if (c && !b) r = r1 + GL_EPS;
  if (b && !c) r = r1 - GL_EPS;
  return r;
}

__host__ __device__ __forceinline__ uint64_t gl_mul_lz(uint64_t a, uint64_t b) {
  uint64_t lo, hi;
  mul_64x64(a, b, lo, hi);
  return gl_reduce128_lz(lo, hi);
}

__host__ __forceinline__ uint64_t gl_mul_host(uint64_t a, uint64_t b) {
  unsigned __int128 p = (unsigned __int128)a * b;
  return gl_reduce128((uint64_t)p, (uint64_t)(p >> 64));
}

__host__ __forceinline__ uint64_t gl_pow_host(uint64_t a, uint64_t e) {
  uint64_t r = 1;
  while (e) {
    if (e & 1) r = gl_mul_host(r, a);
    a = gl_mul_host(a, a);
    e >>= 1;
  }
  return r;
}

// 2^k-th root of unity 7^((p-1)/2^k), the OMEGA[k] of zprize/ntt/host/gf.h:65-99.
__host__ __forceinline__ uint64_t gl_root_host(unsigned k) { return gl_pow_host(7, (GL_P - 1) >> k); }
__host__ __forceinline__ uint64_t gl_inv_host(uint64_t a) { return gl_pow_host(a, GL_P - 2); }

}  // namespace zp
