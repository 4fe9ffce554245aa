// BLS12-377 scalar field Fr (r, 253 bits) in eight 32-bit limbs: the one
// operation the harness's data files need, s -> s 2^256 mod r.
//
// The ZPrize harness deserializes Fr values with ark-ff 0.3 (stored as
// s R_r, R_r = 2^256 mod r) and transmutes the slice to BigInteger256
// (zprize/msm_pippenger/test_fpga_harness/src/util.rs:115-118), so the
// integers an MSM over a TEST_LOAD_DATA_FROM directory multiplies by are
// s 2^256 mod r.  fr_to_montgomery computes that as the Montgomery product
// s * (2^512 mod r) * 2^-256 (CIOS, radix 2^32; r = 1 mod 2^32, so the
// per-row factor is the negated low word).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zp {

struct Fr8 {
  uint32_t w[8];
};

__host__ __device__ __forceinline__ Fr8 fr_modulus() {
  return Fr8{{0x00000001u, 0x0a118000u, 0xd0000001u, 0x59aa76feu, 0x5c37b001u, 0x60b44d1eu, 0x9a2ca556u,
              0x12ab655eu}};
}

// 2^512 mod r
__host__ __device__ __forceinline__ Fr8 fr_r2() {
  return Fr8{{0xb861857bu, 0x25d577bau, 0x8860591fu, 0xcc2c27b5u, 0xe5dc8593u, 0xa7cc008fu, 0xeff1c939u,
              0x011fdae7u}};
}

// a b 2^-256 mod r for a, b < r; result < r.
__host__ __device__ __forceinline__ Fr8 fr_mont_mul(const Fr8 &a, const Fr8 &b) {
  const Fr8 n = fr_modulus();
  uint32_t t[10] = {0};
  for (int i = 0; i < 8; i++) {
    uint64_t c = 0;
    for (int j = 0; j < 8; j++) {
      const uint64_t v = (uint64_t)t[j] + (uint64_t)a.w[j] * b.w[i] + c;
      t[j] = (uint32_t)v;
      c = v >> 32;
    }
    uint64_t v = (uint64_t)t[8] + c;
    t[8] = (uint32_t)v;
    t[9] = (uint32_t)(v >> 32);
    const uint32_t m = 0u - t[0];  // -t0 / r mod 2^32 with r = 1 mod 2^32
    c = ((uint64_t)t[0] + (uint64_t)m * n.w[0]) >> 32;
    for (int j = 1; j < 8; j++) {
      v = (uint64_t)t[j] + (uint64_t)m * n.w[j] + c;
      t[j - 1] = (uint32_t)v;
      c = v >> 32;
    }
    v = (uint64_t)t[8] + c;
    t[7] = (uint32_t)v;
    t[8] = t[9] + (uint32_t)(v >> 32);
  }
  // t < 2r: subtract r once if t >= r.
  Fr8 d;
  uint64_t borrow = 0;
  for (int j = 0; j < 8; j++) {
    const uint64_t v = (uint64_t)t[j] - n.w[j] - borrow;
    d.w[j] = (uint32_t)v;
    borrow = (v >> 63) & 1;
  }
  const bool ge = t[8] != 0 || borrow == 0;
  Fr8 out;
  for (int j = 0; j < 8; j++) out.w[j] = ge ? d.w[j] : t[j];
  return out;
}

__host__ __device__ __forceinline__ Fr8 fr_to_montgomery(const Fr8 &s) { return fr_mont_mul(s, fr_r2()); }

}  // namespace zp
