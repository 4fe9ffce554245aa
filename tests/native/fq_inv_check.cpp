// Host check of fq377.hpp's f_inv_fast (batched binary GCD, f_inv_pornin, plus
// one multiply by R'^3) against f_inv (a^(q-2) by square-and-multiply) and the
// defining property a * a^-1 = 1, on edge values and 200,000 random field
// elements; f_inv_pornin must converge within its step bound on every one.
#include <cstdio>
#include <cstdlib>
#include <random>

#include "msm_host.hpp"

using namespace zp;

int main(int argc, char **argv) {
  const int count = argc > 1 ? atoi(argv[1]) : 20000;
  const FConsts K = make_fconsts();
  std::mt19937_64 rng(377);
  int bad = 0, n = 0;
  auto check = [&](const F &a) {
    F r;
    if (!f_inv_pornin(a, K, r)) bad++;  // the step bound must always suffice
    if (!f_eq(r, f_inv_fast(a, K))) bad++;
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
  // plain 2^k and 2^k - 1 (long runs of zero or one bits: the most and the
  // fewest halvings per subtraction), and (q - 1) / 2^k
  for (int k = 1; k < 377; k++) {
    U384 p2 = {{0, 0, 0, 0, 0, 0}};
    p2.w[k / 64] = 1ull << (k % 64);
    check(f_from_u384(p2));
    U384 m1 = {{0, 0, 0, 0, 0, 0}};
    for (int b = 0; b < k; b++) m1.w[b / 64] |= 1ull << (b % 64);
    check(f_from_u384(m1));
    U384 h = Q_PLAIN;
    h.w[0] -= 1;
    for (int b = 0; b < k; b++) {
      for (int j = 0; j < 5; j++) h.w[j] = (h.w[j] >> 1) | (h.w[j + 1] << 63);
      h.w[5] >>= 1;
    }
    check(f_from_u384(h));
  }
  for (int i = 0; i < count; i++) {
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
