// Host check of fr253.hpp: fr_to_montgomery(s) == s 2^256 mod r against
// 128-bit-limb schoolbook arithmetic, for edge values and random s < r.
#include <cstdio>
#include <random>

#include "fr253.hpp"

using namespace zp;
typedef unsigned __int128 u128;

// Plain big-integer reference: (s << 256) mod r by shift-and-subtract.
static void ref_shift_mod(const uint32_t s[8], uint32_t out[8]) {
  const Fr8 n = fr_modulus();
  uint32_t x[9] = {0};
  for (int i = 0; i < 8; i++) x[i] = s[i];
  for (int bit = 0; bit < 256; bit++) {
    // x = 2x, then subtract r if x >= r (x < r before, so 2x < 2r)
    uint32_t carry = 0;
    for (int i = 0; i < 9; i++) {
      const uint32_t nc = x[i] >> 31;
      x[i] = (x[i] << 1) | carry;
      carry = nc;
    }
    bool ge = x[8] != 0;
    if (!ge) {
      ge = true;
      for (int i = 7; i >= 0; i--)
        if (x[i] != n.w[i]) {
          ge = x[i] > n.w[i];
          break;
        }
    }
    if (ge) {
      int64_t borrow = 0;
      for (int i = 0; i < 9; i++) {
        const int64_t v = (int64_t)x[i] - (i < 8 ? n.w[i] : 0) - borrow;
        x[i] = (uint32_t)v;
        borrow = v < 0;
      }
    }
  }
  for (int i = 0; i < 8; i++) out[i] = x[i];
}

int main() {
  std::mt19937_64 rng(7);
  const Fr8 n = fr_modulus();
  int fails = 0;
  for (int t = 0; t < 20000; t++) {
    Fr8 s{};
    if (t == 0) {
    } else if (t == 1) {
      s.w[0] = 1;
    } else if (t == 2) {
      s = n;
      s.w[0] -= 1;  // r - 1
    } else {
      for (int i = 0; i < 8; i++) s.w[i] = (uint32_t)rng();
      s.w[7] &= 0x0fffffffu;  // < 2^252 < r
    }
    uint32_t want[8];
    ref_shift_mod(s.w, want);
    const Fr8 got = fr_to_montgomery(s);
    for (int i = 0; i < 8; i++)
      if (got.w[i] != want[i]) {
        if (fails++ < 5) printf("FAIL t=%d limb %d\n", t, i);
        break;
      }
  }
  if (fails) {
    printf("%d failures\n", fails);
    return 1;
  }
  printf("ok\n");
  return 0;
}
