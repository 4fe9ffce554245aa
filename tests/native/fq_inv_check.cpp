// Host check of fq377.hpp's f_inv_fast (binary extended Euclid + one multiply
// by R'^3) against f_inv (a^(q-2) by square-and-multiply) and against the
// defining property a * a^-1 = 1, on edge values and random field elements.
#include <cstdio>
#include <random>

#include "msm_host.hpp"

using namespace zp;

int main() {
  const FConsts K = make_fconsts();
  std::mt19937_64 rng(377);
  int bad = 0, n = 0;
  auto check = [&](const F &a) {
    const F r = f_inv_fast(a, K);
    if (f_is_zero_canon(a)) {
      if (!f_is_zero_canon(r)) bad++;
      n++;
      return;
    }
    const F want = f_inv(a, K.one);
    if (!f_eq(r, want)) bad++;
    if (!f_eq(f_mul(a, r), K.one)) bad++;
    n++;
  };
  // edges: 0, 1 (R'), -1, small plain integers, q - small, powers of two
  check(f_zero());
  check(K.one);
  check(f_sub_kq(f_zero(), K.one, K.q2));
  for (int k = 1; k < 40; k++) {
    U384 v = {{(uint64_t)k, 0, 0, 0, 0, 0}};
    check(f_from_u384(v));  // raw limbs: the field element k / R'
    U384 w = Q_PLAIN;
    w.w[0] -= (uint64_t)k;
    check(f_from_u384(w));
    check(f_from_u384(pow2_mod_q(k * 9)));
  }
  for (int i = 0; i < 3000; i++) {
    U384 v;
    for (int j = 0; j < 6; j++) v.w[j] = rng();
    v.w[5] &= (1ull << 57) - 1;  // < 2^377
    while (u384_geq(v, Q_PLAIN)) u384_sub(v, Q_PLAIN);
    check(f_from_u384(v));
  }
  // A planted slip: an inverse off by one multiplication by 2 must fail.
  const F a = f_from_u384(pow2_mod_q(101));
  if (f_eq(f_mul(a, f_add(f_inv_fast(a, K), f_inv_fast(a, K))), K.one)) bad++;
  printf("%d values, %d mismatches\n", n, bad);
  if (bad) return 1;
  printf("ok\n");
  return 0;
}
