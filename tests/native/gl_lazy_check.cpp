// Host-side check of the Goldilocks arithmetic that the 4096-point NTT
// kernels run on the GPU (gl64.hpp, ntt4096.hpp): the same inline functions,
// compiled for the CPU, against plain 128-bit modular arithmetic.
//   - gl_add_lz / gl_sub_lz for any u and canonical t, including t = p - 1;
//   - gl_shl_lz<E> for every E in [1, 96), result canonical and equal to v 2^E;
//   - gl_reduce128_lz / gl_mul_lz for any 64-bit operands;
//   - dft64_regs<0/1> against the defining sum with w_64 = 2^39 (and 2^153),
//     with lazy (>= p) values wherever the precondition allows them, and
//     anywhere when the inputs are canonicalized inside (CANON_IN).
// Prints "ok" and exits 0 when every case passes.
#include <cstdio>
#include <random>
#include <utility>
#include <vector>

#include "ntt4096.hpp"

using namespace zp;
typedef unsigned __int128 u128;

static uint64_t mulmod(uint64_t a, uint64_t b) { return (uint64_t)(((u128)a * b) % GL_P); }
static uint64_t powmod(uint64_t a, uint64_t e) {
  uint64_t r = 1;
  for (; e; e >>= 1, a = mulmod(a, a))
    if (e & 1) r = mulmod(r, a);
  return r;
}

static int fails = 0;
#define EXPECT(cond, what, a, b)                                                                   \
  do {                                                                                             \
    if (!(cond) && fails++ < 20)                                                                   \
      printf("FAIL %s %016llx %016llx\n", what, (unsigned long long)(a), (unsigned long long)(b)); \
  } while (0)

template <int E>
static void check_shl(const std::vector<uint64_t> &vals) {
  const uint64_t w = powmod(2, E);
  for (uint64_t v : vals) {
    const uint64_t r = gl_shl_lz<E>(v);
    EXPECT(r < GL_P && r == mulmod(v % GL_P, w), "shl", v, E);
  }
}
template <int... E>
static void check_all_shl(const std::vector<uint64_t> &vals, std::integer_sequence<int, E...>) {
  (check_shl<E + 1>(vals), ...);
}

int main() {
  std::mt19937_64 rng(2024);
  std::vector<uint64_t> vals = {0, 1, 2, GL_P - 1, GL_P, GL_P + 1, ~0ull, ~0ull - 1, 0xFFFFFFFFull, 0x100000000ull,
                                0xFFFFFFFF00000000ull, 0x8000000000000000ull, 0x7FFFFFFFFFFFFFFFull};
  for (int i = 0; i < 64; i++) {
    const uint64_t b = 1ull << i;
    for (uint64_t v : {b, ~b, b - 1, GL_P - b, (uint64_t)0 - b}) vals.push_back(v);
  }
  for (int i = 0; i < 20000; i++) vals.push_back(rng());
  for (int i = 0; i < 4000; i++) vals.push_back(~0ull - (rng() & 0xFFFFFFFFFull));
  for (int i = 0; i < 4000; i++) vals.push_back(GL_P + (rng() % GL_EPS));

  const size_t nv = vals.size();
  for (size_t i = 0; i < nv; i++) {
    const uint64_t u = vals[i];
    uint64_t t = vals[(i * 7919 + 13) % nv] % GL_P;
    if (i % 3 == 0) t = GL_P - 1 - (rng() & 0xFFFF);
    const uint64_t uu = u % GL_P;
    EXPECT(gl_add_lz(u, t) % GL_P == (uint64_t)(((u128)uu + t) % GL_P), "add", u, t);
    EXPECT(gl_sub_lz(u, t) % GL_P == (uint64_t)(((u128)uu + GL_P - t) % GL_P), "sub", u, t);
    const uint64_t hi = vals[(i * 31 + 7) % nv];
    EXPECT(gl_reduce128_lz(u, hi) % GL_P == (uint64_t)((((u128)hi << 64) | u) % GL_P), "reduce128", u, hi);
    EXPECT(gl_mul_lz(u, hi) % GL_P == (uint64_t)(((u128)u * hi) % GL_P), "mul", u, hi);
    EXPECT(gl_canon(u) == uu, "canon", u, 0);
  }
  check_all_shl(vals, std::make_integer_sequence<int, 95>{});

  for (int dir = 0; dir < 2; dir++) {
    const uint64_t w = dir ? powmod(powmod(2, 39), GL_P - 2) : powmod(2, 39);
    for (int trial = 0; trial < 400; trial++) {
      uint64_t x[64], ref[64];
      for (int i = 0; i < 64; i++) {
        uint64_t v = rng() % GL_P;
        const int mode = (trial + i) % 6;
        if (mode == 0) v = GL_P - 1 - (rng() & 3);
        if (mode == 1) v = rng() & 7;
        if (mode == 2 && i < 32) v = ~0ull - (rng() & 0xFFFFFFFF);  // lazy entries allowed below 32
        if (trial == 0) v = i < 32 ? ~0ull : GL_P - 1;
        if (trial == 1) v = i < 32 ? GL_P : 0;
        x[i] = v;
        ref[i] = v % GL_P;
      }
      // odd trials: any u64 in every slot, canonicalized inside (CANON_IN)
      if (trial & 1)
        for (int i = 32; i < 64; i++)
          if ((trial + i) % 3 == 0) x[i] = GL_P + (rng() % GL_EPS), ref[i] = x[i] % GL_P;
      if (dir)
        (trial & 1) ? dft64_regs<1, true>(x) : dft64_regs<1>(x);
      else
        (trial & 1) ? dft64_regs<0, true>(x) : dft64_regs<0>(x);
      for (int k = 0; k < 64; k++) {
        uint64_t acc = 0;
        for (int m = 0; m < 64; m++) acc = (uint64_t)(((u128)acc + mulmod(ref[m], powmod(w, (uint64_t)m * k))) % GL_P);
        EXPECT(x[k] % GL_P == acc, dir ? "dft64 inverse" : "dft64 forward", x[k], acc);
      }
    }
  }
  if (fails) {
    printf("%d failures\n", fails);
    return 1;
  }
  printf("ok\n");
  return 0;
}
