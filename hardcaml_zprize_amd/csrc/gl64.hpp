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
__device__ __forceinline__ void mul_64x64(uint64_t a, uint64_t b, uint64_t &lo, uint64_t &hi) {
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
