// Host check of msm_plan.hpp (option strings, window layouts, cost model),
// built on its own so that it runs under AddressSanitizer and
// UndefinedBehaviorSanitizer: parse_options reads option strings by length
// (the harness's XCLBIN bytes are not NUL terminated), so it is also fed
// random byte strings cut at random lengths.
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "msm_plan.hpp"

namespace zp {
void set_last_error(const char *, ...) {}
}  // namespace zp

using namespace zp;

static int bad = 0;
#define EXPECT(c)                                         \
  do {                                                    \
    if (!(c)) {                                           \
      printf("failed: %s (line %d)\n", #c, __LINE__);     \
      bad++;                                              \
    }                                                     \
  } while (0)

static MsmOptions parse(const char *s) { return parse_options(s, strlen(s)); }

int main(int argc, char **argv) {
  const int fuzz = argc > 1 ? atoi(argv[1]) : 200000;
  // layouts: every window layout covers exactly 253 bits
  HostLayout L = layout_reference();
  layout_finish(L);
  int bits = 0;
  for (int w = 0; w < L.W; w++) bits += L.width[w];
  EXPECT(L.W == 20 && bits == 253 && L.off[19] == 241 && L.nb[0] == 4096 && L.nb[13] == 2048 && L.nb[19] == 4096);
  for (int c = 8; c <= 24; c++) {
    HostLayout U = layout_uniform(c);
    layout_finish(U);
    int b = 0;
    for (int w = 0; w < U.W; w++) b += U.width[w];
    EXPECT(b == 253 && U.W == (253 + c - 1) / c && U.W <= MSM_MAXW);
    EXPECT(U.nb[U.W - 1] == 1u << (253 - U.off[U.W - 1]));
    for (int w = 0; w + 1 < U.W; w++) EXPECT(U.nb[w] == 1u << (c - 1));
  }
  for (int lg = 0; lg <= 32; lg++) {
    const int c = auto_window(size_t(1) << lg);
    EXPECT(c >= 8 && c <= 20);
  }
  EXPECT(auto_window(0) == 8 && auto_window(1) == 8);
  // option strings
  EXPECT(!parse("").window && !parse("").reference && parse("").precompute == 0);
  EXPECT(parse_options(nullptr, 5).window == 0);
  EXPECT(parse("window=16").window == 16 && parse("window=reference").reference);
  MsmOptions o = parse("precompute=auto,window=12;sort=msd sort_lbits=3");
  EXPECT(o.precompute == -1 && o.window == 12 && o.sort == 2 && o.sort_lbits == 3);
  EXPECT(parse("precompute=yes").bad_precompute && parse("precompute=7").bad_precompute);
  EXPECT(!parse("precompute=8").bad_precompute && !parse("precompute=off").bad_precompute);
  EXPECT(parse("sort=radix").bad_sort && parse("coords=cartesian").bad_coords);
  EXPECT(parse("finalize=cpu").bad_finalize && parse("finalize=host").host_finalize);
  EXPECT(validate_options(parse("window=7"), "t") == ZP_ERR_INVALID_ARG);
  EXPECT(validate_options(parse("window=25"), "t") == ZP_ERR_INVALID_ARG);
  EXPECT(validate_options(parse("window=24,precompute=auto"), "t") == ZP_OK);
  // a path-like XCLBIN string is not an option: defaults
  o = parse("/opt/xilinx/zprize/msm.xclbin");
  EXPECT(o.window == 0 && !o.reference && o.precompute == 0 && validate_options(o, "t") == ZP_OK);
  // the length bounds the string: "window=16" cut to "window=1"
  EXPECT(parse_options("window=16", 8).window == 1);
  EXPECT(parse_options("window=reference,precompute=auto", 16).reference);
  EXPECT(parse_options("window=reference,precompute=auto", 16).precompute == 0);
  // cost model: more points cost more, and the shared bucket set is counted once
  HostLayout A = layout_uniform(16);
  layout_finish(A);
  EXPECT(msm_cost(1000, A, false) < msm_cost(2000, A, false));
  EXPECT(msm_cost(1000, A, true) < msm_cost(1000, A, false));
  // fuzz: random bytes (NUL included) at random lengths, from a heap buffer
  // of exactly that length so that any read past it is caught
  std::mt19937 rng(12345);
  const char alphabet[] = "window=precompute,;auto reference sort msd lds atomic0123456789=,;\\0xyz";
  for (int i = 0; i < fuzz; i++) {
    const size_t len = rng() % 48;
    std::vector<char> buf(len ? len : 1);
    for (size_t k = 0; k < len; k++) buf[k] = alphabet[rng() % (sizeof(alphabet) - 1)];
    const MsmOptions f = parse_options(len ? buf.data() : nullptr, len);
    if (validate_options(f, "fuzz") == ZP_OK) {
      HostLayout C = choose_layout(f, size_t(1) << (rng() % 27));
      int b = 0;
      for (int w = 0; w < C.W; w++) b += C.width[w];
      EXPECT(b == 253 && C.W >= 1 && C.W <= MSM_MAXW);
    }
  }
  printf("%d failures\n", bad);
  if (bad) return 1;
  printf("ok\n");
  return 0;
}
