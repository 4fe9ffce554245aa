/* The oracle's C code under AddressSanitizer and UndefinedBehaviorSanitizer:
 * every entry point run on small inputs and checked against an independent
 * property (brute-force DFT, round trip, four-step = direct, MSM = sum of
 * scalar multiples, digit reconstruction), so that an out-of-bounds access
 * or undefined operation in the checker itself shows up. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static int bad = 0;
#define EXPECT(c)                                        \
  do {                                                   \
    if (!(c)) {                                          \
      printf("failed: %s (line %d)\n", #c, __LINE__);    \
      bad++;                                             \
    }                                                    \
  } while (0)

static uint64_t rng_state = 88172645463325252ull;
static uint64_t rnd(void) {
  rng_state ^= rng_state << 13;
  rng_state ^= rng_state >> 7;
  rng_state ^= rng_state << 17;
  return rng_state;
}

#define P 0xFFFFFFFF00000001ull

int main(void) {
  /* NTT: brute-force DFT for n <= 64, round trips, four-step splits */
  for (unsigned lg = 0; lg <= 10; lg++) {
    const size_t n = (size_t)1 << lg;
    uint64_t *a = malloc(n * 8), *b = malloc(n * 8), *c = malloc(n * 8);
    for (size_t i = 0; i < n; i++) a[i] = rnd() % P;
    memcpy(b, a, n * 8);
    orc_ntt_inplace(b, lg, 0);
    if (lg <= 6) {
      const uint64_t w = orc_gl_root(lg);
      for (size_t k = 0; k < n; k++) {
        uint64_t s = 0;
        for (size_t j = 0; j < n; j++) s = orc_gl_add(s, orc_gl_mul(a[j], orc_gl_pow(w, (uint64_t)(j * k))));
        EXPECT(s == b[k]);
      }
    }
    memcpy(c, b, n * 8);
    orc_ntt_inplace(c, lg, 1);
    EXPECT(memcmp(a, c, n * 8) == 0);
    for (unsigned rows = 0; rows <= lg; rows++) {
      memcpy(c, a, n * 8);
      orc_ntt_four_step(c, lg, rows);
      EXPECT(memcmp(b, c, n * 8) == 0);
    }
    memcpy(c, a, n * 8);
    orc_ntt_many(c, lg, 1, 0, 2);
    EXPECT(memcmp(b, c, n * 8) == 0);
    free(a);
    free(b);
    free(c);
  }
  {
    uint64_t src[256], dst[256];
    for (int i = 0; i < 256; i++) src[i] = rnd() % P;
    orc_ntt_reverse(dst, src, 4);
  }
  /* MSM: digits reconstruct; orc_msm equals the sum of scalar multiples */
  for (int t = 0; t < 2000; t++) {
    orc_bigint256 s;
    for (int w = 0; w < 4; w++) s.data[w] = rnd();
    s.data[3] &= (1ull << 60) - 1;
    uint32_t mag[20];
    uint8_t neg[20];
    orc_msm_digits(&s, mag, neg);
    /* reconstruct with 128-bit words: value = sum +-mag << off */
    unsigned __int128 lo = 0, hi = 0; /* bits 0..127, 128..255 (two's complement over 256 bits) */
    for (int w = 0; w < 20; w++) {
      const int off = orc_msm_window_offset(w);
      unsigned __int128 vlo = 0, vhi = 0;
      const unsigned __int128 m = mag[w];
      if (off < 128) {
        vlo = m << off;
        vhi = off ? m >> (128 - off) : 0;
      } else {
        vhi = m << (off - 128);
      }
      if (neg[w]) { /* subtract */
        const unsigned __int128 nlo = lo - vlo;
        hi = hi - vhi - (nlo > lo);
        lo = nlo;
      } else {
        const unsigned __int128 nlo = lo + vlo;
        hi = hi + vhi + (nlo < lo);
        lo = nlo;
      }
    }
    const unsigned __int128 wl = ((unsigned __int128)s.data[1] << 64) | s.data[0];
    const unsigned __int128 wh = ((unsigned __int128)s.data[3] << 64) | s.data[2];
    EXPECT(lo == wl && hi == wh);
  }
  {
    enum { N = 24 };
    orc_bigint256 ks[N], sc[2 * N];
    orc_g1_affine pts[N];
    for (int i = 0; i < N; i++) {
      for (int w = 0; w < 4; w++) ks[i].data[w] = rnd();
      ks[i].data[3] &= (1ull << 60) - 1;
    }
    for (int i = 0; i < 2 * N; i++) {
      for (int w = 0; w < 4; w++) sc[i].data[w] = rnd();
      sc[i].data[3] &= (1ull << 60) - 1;
    }
    orc_g1_mul_generator_many(pts, ks, N, 2);
    orc_g1_projective out[2];
    orc_msm(pts, N, sc, 2, out, 2);
    for (int b = 0; b < 2; b++) {
      orc_g1_affine acc, t, got;
      memset(&acc, 0, sizeof(acc));
      acc.infinity = 1;
      for (int i = 0; i < N; i++) {
        orc_g1_scalar_mul(&t, &pts[i], &sc[b * N + i]);
        orc_g1_add(&acc, &acc, &t);
      }
      orc_g1_projective_to_affine(&got, &out[b]);
      EXPECT(got.infinity == acc.infinity && (got.infinity || (memcmp(&got.x, &acc.x, 48) == 0 && memcmp(&got.y, &acc.y, 48) == 0)));
    }
    uint64_t d[4];
    orc_dot_mod_r(sc, ks, N, d);
    uint64_t pre[18];
    EXPECT(orc_msm_precompute_point(&pts[0], pre) == 1);
  }
  printf("%d failures\n", bad);
  if (bad) return 1;
  printf("ok\n");
  return 0;
}
