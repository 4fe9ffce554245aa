// Host-side MSM planning: the window layouts, the option string of msm_init
// and the batch calls, and the bucket-addition cost model.  No device code;
// compiled into the library and, for the sanitizer check, on its own
// (tests/native/host_plan_check.cpp).
#pragma once
#include <stdint.h>

#include <algorithm>
#include <cstdlib>
#include <string>

#include "common.hpp"

namespace zp {

// ------------------------------------------------------------ layouts --
static constexpr int MSM_MAXW = 32;

struct HostLayout {
  int W = 0;
  int width[MSM_MAXW] = {};
  int off[MSM_MAXW] = {};
  uint32_t nb[MSM_MAXW] = {};
};

// The reference's 20 windows: 13 of 13 bits then 7 of 12 bits, top window
// unsigned (zprize/msm_pippenger/hardcaml/src/config_utils.ml:7-41).
static HostLayout layout_reference() {
  HostLayout L;
  L.W = 20;
  int off = 0;
  for (int w = 0; w < 20; w++) {
    L.width[w] = w < 13 ? 13 : 12;
    L.off[w] = off;
    off += L.width[w];
  }
  return L;
}

// Uniform windows of c bits (top window takes what is left of the 253 bits).
static HostLayout layout_uniform(int c) {
  HostLayout L;
  L.W = (253 + c - 1) / c;
  for (int w = 0; w < L.W; w++) {
    L.off[w] = w * c;
    L.width[w] = std::min(c, 253 - w * c);
  }
  return L;
}

static void layout_finish(HostLayout &L) {
  for (int w = 0; w < L.W; w++) {
    if (w == L.W - 1)
      L.nb[w] = 1u << (253 - L.off[w]);  // unsigned top window: slice + carry
    else
      L.nb[w] = 1u << (L.width[w] - 1);
  }
}

// Window width for n points (the size-class policy; "window=<c>" overrides).
// c = log2 N - 6 in [8, 20]: measured best or within 7 % at 2^22 (c 14-18:
// 15.5-16.7 ms), 2^24 (c 18: 45 ms; 16: 49-52, 20: 48) and 2^26 (c 20:
// 150-154 ms; 16: 168-171, 22: 164-170).  Windows wider than 15 bits take the
// two-level bucket sort.
static int auto_window(size_t n) {
  int lg = 0;
  while ((1ull << (lg + 1)) <= n) lg++;
  return std::min(20, std::max(8, lg - 6));
}

struct MsmOptions {
  int window = 0;  // 0 = auto
  bool reference = false;
  // msm_init only: point coordinates are plain integers < q ("coords=canonical",
  // the form ark-serialize writes) instead of the wire's Montgomery x R.
  bool canonical_coords = false;
  bool bad_coords = false;  // coords= neither "canonical" nor "montgomery"
  int log_u = -1;  // entries per work unit, -1 = auto
  int log_g = -1;  // buckets per window group, -1 = auto
  // Bucket sort: 0 auto, 1 single-level LDS counting sort (windows of at
  // most 2^15 buckets), 2 two-level (any width), 3 global atomics.
  int sort = 0;
  bool bad_sort = false;
  int sort_lbits = -1;  // two-level: low bucket bits sorted in pass B, -1 = auto
  // msm_init only: also keep the window-shifted table 2^(c w) P_i for the
  // uniform layout of c bits ("precompute=<c>" or "precompute=auto"); 0 = off.
  int precompute = 0;
  bool bad_precompute = false;
  // Batch option "finalize=host": the window combine, the exceptional points
  // (by double-and-add over the exact scalar sums) and the wire conversion
  // run on the host instead of k_msm_finalize (a cross-check of the device
  // path; the results are the same bytes).
  bool host_finalize = false;
  bool bad_finalize = false;
};

static MsmOptions parse_options(const char *s, size_t len) {
  MsmOptions o;
  std::string str = (s && len) ? std::string(s, len) : std::string();
  size_t pos = 0;
  while (pos < str.size()) {
    size_t end = str.find_first_of(", ;", pos);
    if (end == std::string::npos) end = str.size();
    std::string kv = str.substr(pos, end - pos);
    size_t eq = kv.find('=');
    if (eq != std::string::npos) {
      std::string k = kv.substr(0, eq), v = kv.substr(eq + 1);
      if (k == "window") {
        if (v == "reference")
          o.reference = true;
        else
          o.window = atoi(v.c_str());
      } else if (k == "coords") {
        o.canonical_coords = v == "canonical";
        o.bad_coords = v != "canonical" && v != "montgomery";
      } else if (k == "log_u") {
        o.log_u = atoi(v.c_str());
      } else if (k == "log_g") {
        o.log_g = atoi(v.c_str());
      } else if (k == "sort") {
        o.sort = v == "auto" ? 0 : v == "lds" ? 1 : v == "msd" ? 2 : v == "atomic" ? 3 : -1;
        o.bad_sort = o.sort < 0;
      } else if (k == "sort_lbits") {
        o.sort_lbits = atoi(v.c_str());
      } else if (k == "finalize") {
        o.host_finalize = v == "host";
        o.bad_finalize = v != "host" && v != "device";
      } else if (k == "precompute") {
        o.precompute = v == "auto" ? -1 : v == "off" ? 0 : atoi(v.c_str());
        o.bad_precompute = v != "auto" && v != "off" && (o.precompute < 8 || o.precompute > 24);
      }
    }
    pos = end + 1;
  }
  return o;
}

static int validate_options(const MsmOptions &o, const char *who) {
  ZP_CHECK(o.window == 0 || (o.window >= 8 && o.window <= 24), ZP_ERR_INVALID_ARG, "%s: window=%d not in [8, 24]",
           who, o.window);
  ZP_CHECK(o.log_u <= 12 && o.log_g <= 10, ZP_ERR_INVALID_ARG, "%s: log_u/log_g out of range", who);
  ZP_CHECK(!o.bad_coords, ZP_ERR_INVALID_ARG, "%s: coords= must be canonical or montgomery", who);
  ZP_CHECK(!o.bad_sort, ZP_ERR_INVALID_ARG, "%s: sort= must be auto, lds, msd or atomic", who);
  ZP_CHECK(o.sort_lbits <= 12, ZP_ERR_INVALID_ARG, "%s: sort_lbits=%d > 12", who, o.sort_lbits);
  ZP_CHECK(!o.bad_precompute, ZP_ERR_INVALID_ARG, "%s: precompute= must be auto, off or a width in [8, 24]", who);
  ZP_CHECK(!o.bad_finalize, ZP_ERR_INVALID_ARG, "%s: finalize= must be device or host", who);
  return ZP_OK;
}

static HostLayout choose_layout(const MsmOptions &o, size_t n) {
  HostLayout L = o.reference ? layout_reference() : layout_uniform(o.window ? o.window : auto_window(n));
  layout_finish(L);
  return L;
}

// Bucket additions of an n-point job: n per window, plus about 2.6 (full
// additions count ~1.3 mixed ones) per bucket of the running-sum reduction;
// with shared buckets every window adds into the largest window's buckets.
static double msm_cost(uint64_t n, const HostLayout &L, bool shared) {
  double buckets = 0;
  uint32_t maxnb = 0;
  for (int w = 0; w < L.W; w++) {
    buckets += L.nb[w];
    maxnb = std::max(maxnb, L.nb[w]);
  }
  return (double)n * L.W + 2.6 * (shared ? (double)maxnb : buckets);
}

}  // namespace zp
