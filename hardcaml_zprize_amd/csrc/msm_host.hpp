// Host-side big-integer helpers for the MSM constants, and the host part of
// the MSM finalisation (window combine, affine conversion, Weierstrass
// fallback for points without a twisted-Edwards image).  Field arithmetic is
// the same F code the device runs (fq377.hpp, compiled for the host).
#pragma once
#include <stdint.h>

#include <cstring>
#include <string>
#include <vector>

#include "ec.hpp"

namespace zp {

// Plain 384-bit integers (6 little-endian 64-bit words).
struct U384 {
  uint64_t w[6];
};

static const U384 Q_PLAIN = {{0x8508c00000000001ull, 0x170b5d4430000000ull, 0x1ef3622fba094800ull,
                              0x1a22d9f300f5138full, 0xc63b05c06ca1493bull, 0x01ae3a4617c510eaull}};

inline bool u384_geq(const U384 &a, const U384 &b) {
  for (int i = 5; i >= 0; i--) {
    if (a.w[i] != b.w[i]) return a.w[i] > b.w[i];
  }
  return true;
}

inline void u384_sub(U384 &a, const U384 &b) {
  unsigned __int128 br = 0;
  for (int i = 0; i < 6; i++) {
    unsigned __int128 d = (unsigned __int128)a.w[i] - b.w[i] - br;
    a.w[i] = (uint64_t)d;
    br = (d >> 64) & 1;
  }
}

// 2^k mod q by doubling.
inline U384 pow2_mod_q(int k) {
  U384 x = {{1, 0, 0, 0, 0, 0}};
  for (int i = 0; i < k; i++) {
    uint64_t c = 0;
    for (int j = 0; j < 6; j++) {
      uint64_t nc = x.w[j] >> 63;
      x.w[j] = (x.w[j] << 1) | c;
      c = nc;
    }
    if (c || u384_geq(x, Q_PLAIN)) u384_sub(x, Q_PLAIN);
  }
  return x;
}

inline U384 u384_from_hex(const char *hex) {
  U384 r = {{0, 0, 0, 0, 0, 0}};
  size_t n = strlen(hex);
  for (size_t i = 0; i < n; i++) {
    char ch = hex[n - 1 - i];
    uint64_t v = (ch >= '0' && ch <= '9') ? (uint64_t)(ch - '0')
                 : (ch >= 'a' && ch <= 'f') ? (uint64_t)(ch - 'a' + 10)
                                            : (uint64_t)(ch - 'A' + 10);
    r.w[i / 16] |= v << (4 * (i % 16));
  }
  return r;
}

inline F f_from_u384(const U384 &x) {
  Packed p;
  for (int i = 0; i < 6; i++) {
    p.w[2 * i] = (uint32_t)x.w[i];
    p.w[2 * i + 1] = (uint32_t)(x.w[i] >> 32);
  }
  return f_unpack(p);
}

inline U384 u384_from_f(const F &a) {
  Packed p = f_pack(f_canonical(a));
  U384 r;
  for (int i = 0; i < 6; i++) r.w[i] = (uint64_t)p.w[2 * i] | ((uint64_t)p.w[2 * i + 1] << 32);
  return r;
}

// Curve constants of bls12_377_g1.h:31-54 (plain values) and the generator
// of libs/ark_bls12_377_g1/test/test_ark_bls12_377_g1.ml:45-57.
static const char *HEX_D = "e2dcedff103e7161354a88156e4b00fe66a526a0237cfe5f683497c9afb7635d5c9307f78e160e14f2b6d4bd949004";
static const char *HEX_S = "19d47d415b5ff60a87a8b7bbab25eb6427dd58ca38e47030efd1e6310ac7bf3079221bf2b4bd72c5106e9e70fcc6156";
static const char *HEX_TS = "363b01df81a0405e86206f2cc504d147752b91b664d91d4c82fed217902ee874354485836fa3ccba3af539c173857b";
static const char *HEX_CA = "32d756062d349e59416ece15ccbf8e86ef0d33183465a42fe2cb65fc1664272e6bb28f0e1c7a7c9c05824ad09adc01";
static const char *HEX_CB = "19d47d415b5ff60a87a8b7bbab25eb6427dd58ca38e47030efd1e6310ac7bf3079221bf2b4bd72c5106e9e70fcc6156";
static const char *HEX_GX = "8848defe740a67c8fc6225bf87ff5485951e2caa9d41bb188282c8bd37cb5cd5481512ffcd394eeab9b16eb21be9ef";
static const char *HEX_GY = "1914a69c5102eff1f674f5d30afeec4bd7fb348ca3e52d96d182ad44fb82305c2fe3d3634a9591afd82de55559c8ea6";

inline FConsts make_fconsts() {
  FConsts K;
  const F r2 = f_from_u384(pow2_mod_q(784));  // R'^2 mod q
  auto mont = [&](const U384 &plain) { return f_canonical(f_mul(f_from_u384(plain), r2)); };
  K.one = f_canonical(f_from_u384(pow2_mod_q(392)));
  K.two = f_canonical(f_add(K.one, K.one));
  K.to_int = f_from_u384(pow2_mod_q(400));
  K.r2 = r2;
  K.to_wire = f_from_u384(pow2_mod_q(384));
  K.r3 = f_from_u384(pow2_mod_q(3 * 392));
  const F d = mont(u384_from_hex(HEX_D));
  K.d2 = f_canonical(f_add(d, d));
  K.k2d = K.d2;
  K.mont_s = mont(u384_from_hex(HEX_S));
  K.ts = mont(u384_from_hex(HEX_TS));
  U384 qm1 = Q_PLAIN;
  qm1.w[0] -= 1;
  K.alpha = mont(qm1);
  const F ca = mont(u384_from_hex(HEX_CA));
  const F cb = mont(u384_from_hex(HEX_CB));
  K.inv_cb = f_canonical(f_inv(cb, K.one));
  const F three = f_canonical(f_add(K.two, K.one));
  K.ca_3cb = f_canonical(f_mul(ca, f_inv(f_mul(three, cb), K.one)));
  f_make_kq(2, K.q2);
  f_make_kq(4, K.q4);
  f_make_kq(8, K.q8);
  return K;
}

}  // namespace zp
