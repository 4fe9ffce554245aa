// Goldilocks NTT on gfx950: plans, kernels and the NTT half of the C ABI.
//
// A transform of n = 2^k points:
//   k <= 9: one workgroup holds the whole vector in LDS (a single pass);
//   k  > 9: the 4-step of libs/hardcaml_ntt/src/reference_model.ml:92-105 with
//            n = n1 * n2 (by default n1 = 2^ceil(k/2) up to 2^20 and 4096
//            from 2^21; the size-class tile table may choose another split),
//            as the FPGA's two phases (zprize/ntt/host/driver.cpp:201-231):
//     pass 1  column DFTs of length n1 over the n1 x n2 row-major matrix,
//             fused twiddle w_n^(j2 * k1), written to a scratch vector;
//     pass 2  row DFTs of length n2, written transposed into the data vector
//             (natural-order output, n^-1 fused for the inverse).
// HBM traffic is 4 * 8n bytes per transform (two reads, two writes).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "gl64.hpp"
#include "ntt4096.hpp"
#include "ntt4096r16.hpp"
#include "ntt_kernels.hpp"

namespace zp {

static thread_local std::string g_last_error;

void set_last_error(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

// ---------------------------------------------------------------- tables --
// Per-size device tables, shared by every plan and batch job of that size.
struct NttTables {
  unsigned log_n;
  uint64_t *d_sub[2] = {nullptr, nullptr};  // w_4096^i, i < 2048 (dir 0 / dir 1)
  uint64_t *d_lo[2] = {nullptr, nullptr};   // w_n^i, i < 4096
  uint64_t *d_hi[2] = {nullptr, nullptr};   // w_n^(4096 i), i < max(1, n / 4096)
  uint64_t *d_4096[2] = {nullptr, nullptr}; // [k][a] = w_4096^(a k), a, k < 64
  uint64_t *d_full[2] = {nullptr, nullptr}; // w_4096^i, i < 4096
  uint64_t n_inv = 1;
};

static std::mutex g_tables_mu;
static std::map<int, std::map<unsigned, NttTables *>> g_tables;  // device -> log_n -> tables

static int make_tables(unsigned log_n, NttTables **out) {
  int dev = 0;
  ZP_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_tables_mu);
  auto &per_dev = g_tables[dev];
  auto it = per_dev.find(log_n);
  if (it != per_dev.end()) {
    *out = it->second;
    return ZP_OK;
  }
  NttTables *t = new NttTables();
  t->log_n = log_n;
  const uint64_t n = 1ull << log_n;
  const uint64_t n_hi = std::max<uint64_t>(1, n >> 12);
  std::vector<uint64_t> sub(2048), lo(4096), hi(n_hi), full(4096), ka(4096);
  for (int dir = 0; dir < 2; dir++) {
    uint64_t w4096 = gl_root_host(12);
    uint64_t wn = gl_root_host(log_n);
    if (dir) {
      w4096 = gl_inv_host(w4096);
      wn = gl_inv_host(wn);
    }
    uint64_t acc = 1;
    for (int i = 0; i < 4096; i++, acc = gl_mul_host(acc, w4096)) {
      full[i] = acc;
      if (i < 2048) sub[i] = acc;
    }
    acc = 1;
    for (int i = 0; i < 4096; i++, acc = gl_mul_host(acc, wn)) lo[i] = acc;
    uint64_t step = gl_pow_host(wn, 4096);
    acc = 1;
    for (uint64_t i = 0; i < n_hi; i++, acc = gl_mul_host(acc, step)) hi[i] = acc;
    ZP_HIP(hipMalloc(&t->d_sub[dir], 2048 * 8));
    ZP_HIP(hipMalloc(&t->d_4096[dir], 4096 * 8));
    for (int k = 0; k < 64; k++)
      for (int a = 0; a < 64; a++) ka[k * 64 + a] = full[a * k];
    ZP_HIP(hipMemcpy(t->d_4096[dir], ka.data(), 4096 * 8, hipMemcpyHostToDevice));
    ZP_HIP(hipMalloc(&t->d_full[dir], 4096 * 8));
    ZP_HIP(hipMemcpy(t->d_full[dir], full.data(), 4096 * 8, hipMemcpyHostToDevice));
    ZP_HIP(hipMalloc(&t->d_lo[dir], 4096 * 8));
    ZP_HIP(hipMalloc(&t->d_hi[dir], n_hi * 8));
    ZP_HIP(hipMemcpy(t->d_sub[dir], sub.data(), 2048 * 8, hipMemcpyHostToDevice));
    ZP_HIP(hipMemcpy(t->d_lo[dir], lo.data(), 4096 * 8, hipMemcpyHostToDevice));
    ZP_HIP(hipMemcpy(t->d_hi[dir], hi.data(), n_hi * 8, hipMemcpyHostToDevice));
  }
  t->n_inv = gl_inv_host(n % GL_P);
  per_dev[log_n] = t;
  *out = t;
  return ZP_OK;
}

// ------------------------------------------------------------ dispatch --
struct NttShape {
  unsigned log_n, log_n1, log_n2;
  unsigned log_t1 = 0, log_t2 = 0;  // LDS passes: log2 sequences per workgroup
};

// Tile choices per size class (2^10 .. 2^23 points): the column length
// 2^log_n1 (4096: pass 1 on the register-resident column kernel; shorter:
// on the LDS kernels) and the sequences per workgroup of each LDS pass, one
// table for single transforms and one for the groups of a batch (many jobs
// on grid.y).  Fields < 0 take the built-in policy (balanced split up to
// 2^20, 4096-point columns from 2^21).  Set through zp_ntt_set_tiles (the
// autotuner's table); a transform reads its shape once, before its first
// pass.
struct NttTileChoice {
  int log_n1 = -1, log_t1 = -1, log_t2 = -1;
};
static constexpr unsigned TILE_MIN_LOG = 10, TILE_MAX_LOG = 23;
static unsigned default_log_n1(unsigned log_n) { return log_n >= 21 ? 12 : (log_n + 1) / 2; }
static std::mutex g_tile_mu;
static NttTileChoice g_tile_choice[2][TILE_MAX_LOG + 1];

// Up to 2^9 points one workgroup does the whole transform in LDS; above, the
// two passes spread the work over many workgroups.
static NttShape shape_of(unsigned log_n, bool batched = false) {
  NttShape s;
  s.log_n = log_n;
  if (log_n <= 9) {
    s.log_n1 = log_n;
    s.log_n2 = 0;
  } else if (log_n > TILE_MAX_LOG) {
    s.log_n1 = 12;  // 4096 x 4096: register-resident kernels for both passes
    s.log_n2 = log_n - 12;
  } else {
    NttTileChoice c;
    {
      std::lock_guard<std::mutex> lk(g_tile_mu);
      c = g_tile_choice[batched ? 1 : 0][log_n];
    }
    s.log_n1 = c.log_n1 >= 0 ? (unsigned)c.log_n1 : default_log_n1(log_n);
    s.log_n2 = log_n - s.log_n1;
    s.log_t1 = c.log_t1 >= 0 ? (unsigned)c.log_t1 : dft_log_T(s.log_n1, s.log_n2);
    s.log_t2 = c.log_t2 >= 0 ? (unsigned)c.log_t2 : dft_log_T(s.log_n2, s.log_n1);
  }
  return s;
}

// Two-pass transforms keep the 4-step intermediate in an n-element scratch vector.
static bool needs_scratch(unsigned log_n) { return shape_of(log_n).log_n2 > 0; }

// Tile widths (sequences per workgroup) of the 4096-point kernels, per pass:
// ZP_NTT_TILE_COLS / ZP_NTT_TILE_ROWS = 4 or 8 (ZP_NTT_TILE sets both).
static int g_tile_cols = -1, g_tile_rows = -1;

static int env_tile(const char *name, int dflt) {
  const char *e = getenv(name);
  if (!e) e = getenv("ZP_NTT_TILE");
  const int v = e ? atoi(e) : dflt;
  return v == 4 ? 4 : (v == 8 ? 8 : dflt);
}

// Job tables for the batched form: data_tab[j] / scratch_tab[j] (device arrays), jobs = grid.y.
struct JobTabs {
  uint64_t *const *data = nullptr;
  uint64_t *const *scratch = nullptr;
  unsigned jobs = 1;
};

// Pass 1 with 4096-point columns: n2 columns (n2 >= T, a power of two), so
// 2^24 (n2 = 4096) and 2^23 (n2 = 2048) share the register-resident kernel.
template <int T>
static int launch_4096_cols(const NttTables *t, uint64_t *d_in, uint64_t *d_out, int dir, hipStream_t stream,
                            const JobTabs &jt, uint64_t n2 = 4096) {
  Ntt4096Args a{};
  a.n1 = 4096;
  a.n2 = n2;
  a.tw4096 = t->d_4096[dir];
  a.tw_lo = t->d_lo[dir];
  a.tw_hi = t->d_hi[dir];
  a.tiles = (uint32_t)(n2 / T);
  a.in = d_in;
  a.out = d_out;
  a.in_tab = jt.data;
  a.out_tab = jt.scratch;
  a.scale = dir ? t->n_inv : 1;  // applied with the 4-step twiddle; pass 2 does not scale
  const dim3 grid(a.tiles, jt.jobs);
  if (dir)
    hipLaunchKernelGGL((k_ntt4096_cols<T, 1>), grid, dim3(64 * T), ntt4096_lds_bytes<T>(), stream, a);
  else
    hipLaunchKernelGGL((k_ntt4096_cols<T, 0>), grid, dim3(64 * T), ntt4096_lds_bytes<T>(), stream, a);
  ZP_HIP(hipGetLastError());
  return ZP_OK;
}

template <int T>
static int launch_4096_rows(const NttTables *t, uint64_t *d_in, uint64_t *d_out, int dir, hipStream_t stream,
                            const JobTabs &jt) {
  Ntt4096Args a{};
  a.n1 = 4096;
  a.n2 = 4096;
  a.tw4096 = t->d_4096[dir];
  a.tiles = 4096 / T;
  a.in = d_in;
  a.out = d_out;
  a.in_tab = jt.scratch;
  a.out_tab = jt.data;
  a.scale = 1;  // the inverse's n^-1 is applied in pass 1
  const dim3 grid(a.tiles, jt.jobs);
  if (dir)
    hipLaunchKernelGGL((k_ntt4096_rows<T, 1>), grid, dim3(64 * T), ntt4096_lds_bytes<T>(), stream, a);
  else
    hipLaunchKernelGGL((k_ntt4096_rows<T, 0>), grid, dim3(64 * T), ntt4096_lds_bytes<T>(), stream, a);
  ZP_HIP(hipGetLastError());
  return ZP_OK;
}

// Pass 2 of the 2^24 transform: 16 x 16 x 16 row DFTs (ntt4096r16.hpp).
static int launch_r16_rows(const NttTables *t, uint64_t *d_in, uint64_t *d_out, int dir, hipStream_t stream,
                           const JobTabs &jt) {
  Ntt16Args a{};
  a.n1 = 4096;
  a.n2 = 4096;
  a.tw_full = t->d_full[dir];
  a.tiles = 4096 / R16_T;
  a.in = d_in;
  a.out = d_out;
  a.in_tab = jt.scratch;
  a.out_tab = jt.data;
  const dim3 grid(a.tiles, jt.jobs);
  if (dir)
    hipLaunchKernelGGL(k_ntt4096r16_rows<1>, grid, dim3(256 * R16_T), r16_lds_bytes(), stream, a);
  else
    hipLaunchKernelGGL(k_ntt4096r16_rows<0>, grid, dim3(256 * R16_T), r16_lds_bytes(), stream, a);
  ZP_HIP(hipGetLastError());
  return ZP_OK;
}

// ----------------------------------------------------------- REVERSE core --
// The FPGA NTT build's memory-path debug core (zprize/ntt/host/driver.h:18,
// reference.cpp:60-84): the two phases of the 4-step transform with each row
// DFT replaced by reversing the row.  With R = 2^log_r and w = OMEGA[2 log_r]:
//   phase 1 (reference.cpp:69-77)  E[i][j] = in[R-1-i][j] * w^(i j)
//   phase 2 (reference.cpp:79-83)  out[i][j] = E[j][R-1-i]
// (reverse columns + twiddle, then reverse rows + transpose).
__global__ void __launch_bounds__(256) k_reverse_phase1(const uint64_t *in, uint64_t *out, unsigned log_r,
                                                        const uint64_t *tw_lo, const uint64_t *tw_hi) {
  const uint64_t r = 1ull << log_r, n = r << log_r;
  for (uint64_t idx = blockIdx.x * 256ull + threadIdx.x; idx < n; idx += gridDim.x * 256ull) {
    const uint64_t i = idx >> log_r, j = idx & (r - 1);
    const uint64_t e = (i * j) & (n - 1);  // w^n = 1
    const uint64_t w = gl_mul(tw_hi[e >> 12], tw_lo[e & 4095]);
    out[idx] = gl_mul(in[((r - 1 - i) << log_r) + j], w);
  }
}

// 32 x 32 tiles through LDS so that both the reads (rows of E) and the
// writes (rows of out) are contiguous; R < 32 takes the direct gather.
__global__ void __launch_bounds__(256) k_reverse_phase2(const uint64_t *in, uint64_t *out, unsigned log_r) {
  const uint64_t r = 1ull << log_r;
  if (r < 32) {
    const uint64_t idx = threadIdx.x;  // one block of >= r * r threads
    if (idx < r * r) {
      const uint64_t i = idx >> log_r, j = idx & (r - 1);
      out[idx] = in[(j << log_r) + (r - 1 - i)];
    }
    return;
  }
  __shared__ uint64_t tile[32][33];
  const uint32_t tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  const uint64_t c0 = (uint64_t)blockIdx.x * 32, j0 = (uint64_t)blockIdx.y * 32;
  for (int k = 0; k < 32; k += 8) tile[ty + k][tx] = in[((j0 + ty + k) << log_r) + c0 + tx];  // E[j][c]
  __syncthreads();
  // out[R-1-c][j] = E[j][c]
  for (int k = 0; k < 32; k += 8) out[((r - 1 - (c0 + ty + k)) << log_r) + j0 + tx] = tile[tx][ty + k];
}

static int launch_reverse_phase(const NttTables *t, uint64_t *d_data, uint64_t *d_scratch, int pass,
                                hipStream_t stream) {
  const unsigned log_r = t->log_n / 2;
  const uint64_t r = 1ull << log_r, n = r * r;
  if (pass == 1) {
    const unsigned blocks = (unsigned)std::min<uint64_t>((n + 255) / 256, 1u << 20);
    hipLaunchKernelGGL(k_reverse_phase1, dim3(blocks), dim3(256), 0, stream, d_data, d_scratch, log_r,
                       t->d_lo[0], t->d_hi[0]);
  } else if (r < 32) {
    hipLaunchKernelGGL(k_reverse_phase2, dim3(1), dim3(256), 0, stream, d_scratch, d_data, log_r);
  } else {
    hipLaunchKernelGGL(k_reverse_phase2, dim3((unsigned)(r / 32), (unsigned)(r / 32)), dim3(256), 0, stream,
                       d_scratch, d_data, log_r);
  }
  ZP_HIP(hipGetLastError());
  return ZP_OK;
}

static bool g_tile_rows_env = false, g_tile_cols_env = false;
static void tile_widths() {
  if (g_tile_cols < 0) {
    g_tile_cols = env_tile("ZP_NTT_TILE_COLS", 8);
    g_tile_cols_env = getenv("ZP_NTT_TILE_COLS") || getenv("ZP_NTT_TILE");
  }
  if (g_tile_rows < 0) {
    g_tile_rows = env_tile("ZP_NTT_TILE_ROWS", 8);
    g_tile_rows_env = getenv("ZP_NTT_TILE_ROWS") || getenv("ZP_NTT_TILE");
  }
}

// One pass of an in-place transform (pass 1 = column DFTs + 4-step twiddle,
// data -> scratch; pass 2 = row DFTs, scratch -> data transposed).  A length
// that is a single DFT (log_n <= 9) has only pass 1, which works in place.
// scratch must hold n elements otherwise (needs_scratch).
static int launch_pass(const NttTables *t, const NttShape &sh, uint64_t *d_data, uint64_t *d_scratch, int dir,
                       int pass, hipStream_t stream, const JobTabs &jt = JobTabs()) {
  if (sh.log_n1 == 12 && sh.log_n2 == 12) {
    tile_widths();
    if (pass == 1)
      return g_tile_cols == 4 ? launch_4096_cols<4>(t, d_data, d_scratch, dir, stream, jt)
                              : launch_4096_cols<8>(t, d_data, d_scratch, dir, stream, jt);
    // ZP_NTT_TILE_ROWS=4 or 8 selects the 64 x 64 row kernel instead
    if (g_tile_rows_env) return g_tile_rows == 4 ? launch_4096_rows<4>(t, d_scratch, d_data, dir, stream, jt)
                                                 : launch_4096_rows<8>(t, d_scratch, d_data, dir, stream, jt);
    return launch_r16_rows(t, d_scratch, d_data, dir, stream, jt);
  }
  if (pass == 1 && sh.log_n1 == 12 && sh.log_n2 >= 3) {
    // 4096-point columns in registers; pass 2 (rows shorter than 4096) stays
    // on the LDS kernel, which takes the lazy scratch values (canonicalized
    // on load).
    tile_widths();
    const uint64_t n2 = 1ull << sh.log_n2;
    // Fewer 8-wide tiles than CUs (2^21, 2^22 alone): 4-wide tiles, two per
    // CU, fill the GPU (2^21 pass 1 55.9 -> 45.9 us, 2^22 58.6 -> 47.6 us).
    const int tc = g_tile_cols_env ? g_tile_cols : ((n2 >> 3) * jt.jobs < 256 ? 4 : 8);
    return tc == 4 ? launch_4096_cols<4>(t, d_data, d_scratch, dir, stream, jt, n2)
                   : launch_4096_cols<8>(t, d_data, d_scratch, dir, stream, jt, n2);
  }
  const uint64_t n = 1ull << sh.log_n;
  const uint64_t scale = dir ? t->n_inv : 1;
  if (sh.log_n2 == 0) {
    if (pass != 1) return ZP_OK;
    DftArgs a{};
    a.in = d_data;
    a.out = d_data;
    a.log_L = sh.log_n;
    a.log_T = 0;
    a.in_elem_stride = 1;
    a.in_seq_stride = n;
    a.out_elem_stride = 1;
    a.out_seq_stride = n;
    a.tw_sub = t->d_sub[dir];
    a.tw_full = t->d_full[dir];
    a.dir = dir;
    a.scale = scale;
    a.in_tab = jt.data;
    a.out_tab = jt.data;
    return launch_dft(a, 1, stream, jt.jobs);
  }
  const uint64_t n1 = 1ull << sh.log_n1, n2 = 1ull << sh.log_n2;
  if (pass == 1) {
    // Columns (length n1, n2 of them) + twiddle, data -> scratch.
    DftArgs p1{};
    p1.in = d_data;
    p1.out = d_scratch;
    p1.log_L = sh.log_n1;
    p1.log_T = sh.log_t1;
    p1.in_elem_stride = n2;
    p1.in_seq_stride = 1;
    p1.out_elem_stride = n2;
    p1.out_seq_stride = 1;
    p1.tw_sub = t->d_sub[dir];
    p1.tw_full = t->d_full[dir];
    p1.dir = dir;
    p1.tw_lo = t->d_lo[dir];
    p1.tw_hi = t->d_hi[dir];
    p1.scale = 1;
    p1.in_tab = jt.data;
    p1.out_tab = jt.scratch;
    return launch_dft(p1, n2 >> p1.log_T, stream, jt.jobs);
  }
  // Rows (length n2, n1 of them), scratch -> data transposed.
  DftArgs p2{};
  p2.in = d_scratch;
  p2.out = d_data;
  p2.log_L = sh.log_n2;
  p2.log_T = sh.log_t2;
  p2.in_elem_stride = 1;
  p2.in_seq_stride = n2;
  p2.out_elem_stride = n1;
  p2.out_seq_stride = 1;
  p2.tw_sub = t->d_sub[dir];
  p2.tw_full = t->d_full[dir];
  p2.dir = dir;
  // after the register-resident 4096-point columns, n^-1 is already applied
  p2.scale = (sh.log_n1 == 12 && sh.log_n2 >= 3) ? 1 : scale;
  p2.in_tab = jt.scratch;
  p2.out_tab = jt.data;
  return launch_dft(p2, n1 >> p2.log_T, stream, jt.jobs);
}

// Enqueue one in-place transform.
static int launch_transform(const NttTables *t, uint64_t *d_data, uint64_t *d_scratch, int dir,
                            hipStream_t stream, const JobTabs &jt = JobTabs()) {
  const NttShape sh = shape_of(t->log_n, jt.data != nullptr);
  ZP_TRY(launch_pass(t, sh, d_data, d_scratch, dir, 1, stream, jt));
  return launch_pass(t, sh, d_data, d_scratch, dir, 2, stream, jt);
}

}  // namespace zp

using namespace zp;

// ------------------------------------------------------------------ plan --
// A buffer of the asynchronous API: pinned host input/output vectors and a
// device vector.  An evaluation is copy-in (h2d stream) -> transform (the
// plan's stream, which serializes transforms: they share the plan's 4-step
// scratch) -> copy-out (d2h stream), chained by events, so the copies of
// one buffer overlap the transforms of others.
struct UserBuffer {
  uint64_t *h_in = nullptr;
  uint64_t *h_out = nullptr;
  uint64_t *d_buf = nullptr;
  hipEvent_t in_done = nullptr;    // copy-in finished
  hipEvent_t comp_done = nullptr;  // transform finished
  hipEvent_t done = nullptr;       // copy-out finished (the evaluation's end)
  bool in_use = false;
  bool enqueued = false;  // evaluated at least once since the request
};

struct zp_ntt_plan {
  unsigned log_n = 0;
  uint64_t n = 0;
  bool reverse = false;  // the REVERSE debug core instead of the transform
  NttTables *tables = nullptr;
  uint64_t *d_scratch = nullptr;  // n elements (4-step intermediate)
  uint64_t *d_work = nullptr;     // n elements for simple_evaluate
  hipStream_t stream = nullptr;  // transforms (and simple_evaluate)
  hipStream_t h2d = nullptr;     // buffer API copy-in
  hipStream_t d2h = nullptr;     // buffer API copy-out
  std::vector<UserBuffer> buffers;
};

// One pass (1 or 2) of the plan's operation on d_data (in place, scratch in between).
static int plan_pass(const zp_ntt_plan *p, const NttShape &sh, uint64_t *d_data, int dir, int pass,
                     hipStream_t stream) {
  if (p->reverse) return launch_reverse_phase(p->tables, d_data, p->d_scratch, pass, stream);
  return launch_pass(p->tables, sh, d_data, p->d_scratch, dir, pass, stream);
}

static int plan_launch(const zp_ntt_plan *p, uint64_t *d_data, int dir, hipStream_t stream) {
  const NttShape sh = shape_of(p->tables->log_n);
  ZP_TRY(plan_pass(p, sh, d_data, dir, 1, stream));
  return plan_pass(p, sh, d_data, dir, 2, stream);
}

static int check_dir(const zp_ntt_plan *p, int dir, const char *who) {
  ZP_CHECK(dir == 0 || dir == 1, ZP_ERR_INVALID_ARG, "%s: dir %d not 0/1", who, dir);
  ZP_CHECK(!(p->reverse && dir == 1), ZP_ERR_INVALID_ARG, "%s: the REVERSE core has no inverse", who);
  return ZP_OK;
}

extern "C" const char *zp_last_error(void) { return g_last_error.c_str(); }
extern "C" const char *zp_version(void) { return "zpgpu gfx950"; }

extern "C" int zp_ntt_set_tiles(unsigned log_n, int batched, int log_n1, int log_t1, int log_t2) {
  ZP_CHECK(log_n >= TILE_MIN_LOG && log_n <= TILE_MAX_LOG, ZP_ERR_INVALID_ARG,
           "zp_ntt_set_tiles: log_n %u not in [%u, %u]", log_n, TILE_MIN_LOG, TILE_MAX_LOG);
  ZP_CHECK(batched == 0 || batched == 1, ZP_ERR_INVALID_ARG, "zp_ntt_set_tiles: batched %d not 0/1", batched);
  ZP_CHECK(log_n1 >= -1 && log_t1 >= -1 && log_t2 >= -1, ZP_ERR_INVALID_ARG,
           "zp_ntt_set_tiles: fields must be >= -1 (got %d, %d, %d)", log_n1, log_t1, log_t2);
  const int ln = (int)log_n;
  const int n1 = log_n1 >= 0 ? log_n1 : (int)default_log_n1(log_n), n2 = ln - n1;
  ZP_CHECK(n1 >= 1 && n1 <= 12 && n2 >= 1 && n2 <= 12, ZP_ERR_INVALID_ARG,
           "zp_ntt_set_tiles: 2^%d = 2^%d x 2^%d: both sub-DFT lengths must be 2^1 .. 2^12", ln, n1, n2);
  ZP_CHECK(log_t1 <= std::min(13 - n1, n2), ZP_ERR_INVALID_ARG,
           "zp_ntt_set_tiles: log_t1 %d > min(13 - %d, %d) (LDS holds 2^13 elements)", log_t1, n1, n2);
  ZP_CHECK(log_t2 <= std::min(13 - n2, n1), ZP_ERR_INVALID_ARG,
           "zp_ntt_set_tiles: log_t2 %d > min(13 - %d, %d) (LDS holds 2^13 elements)", log_t2, n2, n1);
  std::lock_guard<std::mutex> lk(g_tile_mu);
  g_tile_choice[batched][log_n] = NttTileChoice{log_n1, log_t1, log_t2};
  return ZP_OK;
}

extern "C" int zp_ntt_get_tiles(unsigned log_n, int batched, int *log_n1, int *log_t1, int *log_t2) {
  ZP_CHECK(log_n >= TILE_MIN_LOG && log_n <= TILE_MAX_LOG, ZP_ERR_INVALID_ARG,
           "zp_ntt_get_tiles: log_n %u not in [%u, %u]", log_n, TILE_MIN_LOG, TILE_MAX_LOG);
  ZP_CHECK(batched == 0 || batched == 1, ZP_ERR_INVALID_ARG, "zp_ntt_get_tiles: batched %d not 0/1", batched);
  ZP_CHECK(log_n1 && log_t1 && log_t2, ZP_ERR_INVALID_ARG, "zp_ntt_get_tiles: NULL output");
  const NttShape sh = shape_of(log_n, batched != 0);
  *log_n1 = (int)sh.log_n1;
  *log_t1 = (int)sh.log_t1;
  *log_t2 = (int)sh.log_t2;
  return ZP_OK;
}

static int create_plan(unsigned log_n, unsigned num_buffers, bool reverse, zp_ntt_plan **out, const char *who) {
  ZP_CHECK(out != nullptr, ZP_ERR_INVALID_ARG, "%s: out is NULL", who);
  *out = nullptr;
  if (num_buffers == 0) num_buffers = 8;
  ZP_CHECK(num_buffers <= 64, ZP_ERR_INVALID_ARG, "%s: num_buffers %u > 64", who, num_buffers);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    set_last_error("%s: no HIP device", who);
    return ZP_ERR_NO_DEVICE;
  }
  zp_ntt_plan *p = new zp_ntt_plan();
  p->log_n = log_n;
  p->n = 1ull << log_n;
  p->reverse = reverse;
  int st = make_tables(log_n, &p->tables);
  if (st == ZP_OK && (needs_scratch(log_n) || reverse) && hipMalloc(&p->d_scratch, p->n * 8) != hipSuccess)
    st = ZP_ERR_OUT_OF_MEMORY;
  if (st == ZP_OK && hipMalloc(&p->d_work, p->n * 8) != hipSuccess) st = ZP_ERR_OUT_OF_MEMORY;
  if (st == ZP_OK && hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) st = ZP_ERR_HIP;
  if (st == ZP_OK && hipStreamCreateWithFlags(&p->h2d, hipStreamNonBlocking) != hipSuccess) st = ZP_ERR_HIP;
  if (st == ZP_OK && hipStreamCreateWithFlags(&p->d2h, hipStreamNonBlocking) != hipSuccess) st = ZP_ERR_HIP;
  if (st == ZP_OK) {
    p->buffers.resize(num_buffers);
  }
  if (st != ZP_OK) {
    if (st == ZP_ERR_OUT_OF_MEMORY) set_last_error("%s: device allocation failed", who);
    zp_ntt_destroy(p);
    return st;
  }
  *out = p;
  return ZP_OK;
}

extern "C" int zp_ntt_create(unsigned log_n, unsigned num_buffers, zp_ntt_plan **out) {
  ZP_CHECK(log_n >= 1 && log_n <= 24, ZP_ERR_INVALID_ARG, "zp_ntt_create: log_n %u not in [1, 24]", log_n);
  return create_plan(log_n, num_buffers, false, out, "zp_ntt_create");
}

extern "C" int zp_ntt_create_reverse(unsigned log_row_size, unsigned num_buffers, zp_ntt_plan **out) {
  ZP_CHECK(log_row_size >= 1 && log_row_size <= 12, ZP_ERR_INVALID_ARG,
           "zp_ntt_create_reverse: log_row_size %u not in [1, 12]", log_row_size);
  return create_plan(2 * log_row_size, num_buffers, true, out, "zp_ntt_create_reverse");
}

extern "C" int zp_ntt_destroy(zp_ntt_plan *p) {
  if (!p) return ZP_OK;
  for (hipStream_t st : {p->h2d, p->stream, p->d2h})
    if (st) (void)hipStreamSynchronize(st);
  for (auto &b : p->buffers) {
    if (b.h_in) (void)hipHostFree(b.h_in);
    if (b.h_out) (void)hipHostFree(b.h_out);
    if (b.d_buf) (void)hipFree(b.d_buf);
    for (hipEvent_t e : {b.in_done, b.comp_done, b.done})
      if (e) (void)hipEventDestroy(e);
  }
  if (p->d_scratch) (void)hipFree(p->d_scratch);
  if (p->d_work) (void)hipFree(p->d_work);
  for (hipStream_t st : {p->h2d, p->stream, p->d2h})
    if (st) (void)hipStreamDestroy(st);
  delete p;
  return ZP_OK;
}

extern "C" uint64_t zp_ntt_size(const zp_ntt_plan *p) { return p ? p->n : 0; }
extern "C" unsigned zp_ntt_max_buffers_in_flight(const zp_ntt_plan *p) {
  return p ? (unsigned)p->buffers.size() : 0;
}

extern "C" int zp_ntt_inplace_device(zp_ntt_plan *p, uint64_t *d_data, int dir, void *stream) {
  ZP_CHECK(p != nullptr && d_data != nullptr, ZP_ERR_INVALID_ARG, "zp_ntt_inplace_device: NULL argument");
  ZP_TRY(check_dir(p, dir, "zp_ntt_inplace_device"));
  return plan_launch(p, d_data, dir, (hipStream_t)stream);
}

extern "C" int zp_ntt_time_passes(zp_ntt_plan *p, uint64_t *d_data, int dir, int reps, void *stream, float *ms_pass1,
                                  float *ms_pass2) {
  ZP_CHECK(p != nullptr && d_data != nullptr && ms_pass1 != nullptr && ms_pass2 != nullptr, ZP_ERR_INVALID_ARG,
           "zp_ntt_time_passes: NULL argument");
  ZP_TRY(check_dir(p, dir, "zp_ntt_time_passes"));
  ZP_CHECK(reps > 0, ZP_ERR_INVALID_ARG, "zp_ntt_time_passes: reps %d", reps);
  const hipStream_t s = (hipStream_t)stream;
  std::vector<hipEvent_t> ev(3 * (size_t)reps);
  for (auto &e : ev) ZP_HIP(hipEventCreate(&e));
  int rc = ZP_OK;
  const NttShape sh = shape_of(p->tables->log_n);
  for (int r = 0; r < reps && rc == ZP_OK; r++) {
    (void)hipEventRecord(ev[3 * r], s);
    rc = plan_pass(p, sh, d_data, dir, 1, s);
    (void)hipEventRecord(ev[3 * r + 1], s);
    if (rc == ZP_OK) rc = plan_pass(p, sh, d_data, dir, 2, s);
    (void)hipEventRecord(ev[3 * r + 2], s);
  }
  double t1 = 0, t2 = 0;
  if (rc == ZP_OK && hipEventSynchronize(ev.back()) == hipSuccess) {
    for (int r = 0; r < reps; r++) {
      float a = 0, b = 0;
      (void)hipEventElapsedTime(&a, ev[3 * r], ev[3 * r + 1]);
      (void)hipEventElapsedTime(&b, ev[3 * r + 1], ev[3 * r + 2]);
      t1 += a;
      t2 += b;
    }
  } else if (rc == ZP_OK) {
    set_last_error("zp_ntt_time_passes: event synchronize failed");
    rc = ZP_ERR_HIP;
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  *ms_pass1 = (float)(t1 / reps);
  *ms_pass2 = (float)(t2 / reps);
  return rc;
}

extern "C" int zp_ntt_simple_evaluate(zp_ntt_plan *p, uint64_t *out, const uint64_t *in, uint64_t len, int dir) {
  ZP_CHECK(p != nullptr && out != nullptr && (in != nullptr || len == 0), ZP_ERR_INVALID_ARG,
           "zp_ntt_simple_evaluate: NULL argument");
  ZP_TRY(check_dir(p, dir, "zp_ntt_simple_evaluate"));
  ZP_CHECK(len <= p->n, ZP_ERR_TOO_LARGE, "zp_ntt_simple_evaluate: data_length %llu > %llu",
           (unsigned long long)len, (unsigned long long)p->n);
  if (len) ZP_HIP(hipMemcpyAsync(p->d_work, in, len * 8, hipMemcpyHostToDevice, p->stream));
  if (len < p->n) ZP_HIP(hipMemsetAsync(p->d_work + len, 0, (p->n - len) * 8, p->stream));
  ZP_TRY(plan_launch(p, p->d_work, dir, p->stream));
  ZP_HIP(hipMemcpyAsync(out, p->d_work, p->n * 8, hipMemcpyDeviceToHost, p->stream));
  ZP_HIP(hipStreamSynchronize(p->stream));
  return ZP_OK;
}

static int check_buffer(zp_ntt_plan *p, int id) {
  ZP_CHECK(p != nullptr, ZP_ERR_INVALID_ARG, "NULL plan");
  ZP_CHECK(id >= 0 && id < (int)p->buffers.size(), ZP_ERR_INVALID_ARG, "buffer id %d out of range", id);
  return ZP_OK;
}

extern "C" int zp_ntt_request_buffer(zp_ntt_plan *p, int *id) {
  ZP_CHECK(p != nullptr && id != nullptr, ZP_ERR_INVALID_ARG, "zp_ntt_request_buffer: NULL argument");
  for (size_t i = 0; i < p->buffers.size(); i++) {
    UserBuffer &b = p->buffers[i];
    if (b.in_use) continue;
    if (!b.h_in) {
      ZP_HIP(hipHostMalloc(&b.h_in, p->n * 8, hipHostMallocDefault));
      ZP_HIP(hipHostMalloc(&b.h_out, p->n * 8, hipHostMallocDefault));
      ZP_HIP(hipMalloc(&b.d_buf, p->n * 8));
      ZP_HIP(hipEventCreateWithFlags(&b.in_done, hipEventDisableTiming));
      ZP_HIP(hipEventCreateWithFlags(&b.comp_done, hipEventDisableTiming));
      ZP_HIP(hipEventCreateWithFlags(&b.done, hipEventDisableTiming));
    }
    b.in_use = true;
    b.enqueued = false;
    *id = (int)i;
    return ZP_OK;
  }
  *id = -1;
  set_last_error("zp_ntt_request_buffer: all %zu buffers in flight", p->buffers.size());
  return ZP_ERR_NO_BUFFER;
}

extern "C" uint64_t *zp_ntt_buffer_input(zp_ntt_plan *p, int id) {
  if (check_buffer(p, id) != ZP_OK || !p->buffers[id].in_use) return nullptr;
  return p->buffers[id].h_in;
}

extern "C" uint64_t *zp_ntt_buffer_output(zp_ntt_plan *p, int id) {
  if (check_buffer(p, id) != ZP_OK || !p->buffers[id].in_use) return nullptr;
  return p->buffers[id].h_out;
}

// driver.cpp:277-284.  A buffer still in flight may be enqueued again
// (bench_throughput.cpp:76-86 enqueues each buffer many times before
// waiting): the new copy-in waits for the previous copy-out of the buffer,
// and every evaluation recomputes the output from the pinned input.
extern "C" int zp_ntt_enqueue(zp_ntt_plan *p, int id, int dir) {
  ZP_TRY(check_buffer(p, id));
  UserBuffer &b = p->buffers[id];
  ZP_CHECK(b.in_use, ZP_ERR_STATE, "zp_ntt_enqueue: buffer %d not requested", id);
  ZP_TRY(check_dir(p, dir, "zp_ntt_enqueue"));
  if (b.enqueued) ZP_HIP(hipStreamWaitEvent(p->h2d, b.done, 0));
  ZP_HIP(hipMemcpyAsync(b.d_buf, b.h_in, p->n * 8, hipMemcpyHostToDevice, p->h2d));
  ZP_HIP(hipEventRecord(b.in_done, p->h2d));
  ZP_HIP(hipStreamWaitEvent(p->stream, b.in_done, 0));
  ZP_TRY(plan_launch(p, b.d_buf, dir, p->stream));
  ZP_HIP(hipEventRecord(b.comp_done, p->stream));
  ZP_HIP(hipStreamWaitEvent(p->d2h, b.comp_done, 0));
  ZP_HIP(hipMemcpyAsync(b.h_out, b.d_buf, p->n * 8, hipMemcpyDeviceToHost, p->d2h));
  ZP_HIP(hipEventRecord(b.done, p->d2h));
  b.enqueued = true;
  return ZP_OK;
}

// driver.cpp:246-249: blocks until the buffer's most recent evaluation has
// been copied out.  Waiting again returns at once.
extern "C" int zp_ntt_wait(zp_ntt_plan *p, int id) {
  ZP_TRY(check_buffer(p, id));
  UserBuffer &b = p->buffers[id];
  ZP_CHECK(b.in_use && b.enqueued, ZP_ERR_STATE, "zp_ntt_wait: buffer %d not enqueued", id);
  ZP_HIP(hipEventSynchronize(b.done));
  return ZP_OK;
}

extern "C" int zp_ntt_free_buffer(zp_ntt_plan *p, int id) {
  ZP_TRY(check_buffer(p, id));
  UserBuffer &b = p->buffers[id];
  // driver.cpp:160-170 raises on a double free.
  ZP_CHECK(b.in_use, ZP_ERR_STATE, "zp_ntt_free_buffer: buffer %d is not in use", id);
  if (b.enqueued) {
    ZP_HIP(hipEventSynchronize(b.done));
    b.enqueued = false;
  }
  b.in_use = false;
  return ZP_OK;
}

// ----------------------------------------------------------------- batch --
// Jobs grouped by (log_n, dir): one launch sequence per group with the job
// index on grid.y, per-job data / scratch pointers in device tables.  The
// whole sequence is captured into a HIP graph at the first launch and
// replayed afterwards.
struct NttGroup {
  unsigned log_n;
  int dir;
  unsigned jobs;
  size_t tab_off;  // index of this group's first entry in the pointer tables
};

struct zp_ntt_batch {
  std::vector<NttGroup> groups;
  uint64_t *d_scratch = nullptr;  // arena
  uint64_t **d_tabs = nullptr;    // [data tables | scratch tables]
  size_t count = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

static int batch_launch_eager(zp_ntt_batch *b, hipStream_t stream) {
  for (const NttGroup &g : b->groups) {
    NttTables *t = nullptr;
    ZP_TRY(make_tables(g.log_n, &t));
    JobTabs jt;
    jt.data = b->d_tabs + g.tab_off;
    jt.scratch = b->d_tabs + b->count + g.tab_off;
    jt.jobs = g.jobs;
    ZP_TRY(launch_transform(t, nullptr, nullptr, g.dir, stream, jt));
  }
  return ZP_OK;
}

extern "C" int zp_ntt_batch_create(uint64_t *const *d_ptrs, const unsigned *log_ns, const int *dirs, size_t count,
                                   zp_ntt_batch **out) {
  ZP_CHECK(out != nullptr, ZP_ERR_INVALID_ARG, "zp_ntt_batch_create: out is NULL");
  *out = nullptr;
  ZP_CHECK(count == 0 || (d_ptrs && log_ns && dirs), ZP_ERR_INVALID_ARG, "zp_ntt_batch_create: NULL argument");
  for (size_t i = 0; i < count; i++) {
    ZP_CHECK(log_ns[i] >= 1 && log_ns[i] <= 24, ZP_ERR_INVALID_ARG, "job %zu: log_n %u not in [1, 24]", i, log_ns[i]);
    ZP_CHECK(dirs[i] == 0 || dirs[i] == 1, ZP_ERR_INVALID_ARG, "job %zu: dir %d", i, dirs[i]);
    ZP_CHECK(d_ptrs[i] != nullptr, ZP_ERR_INVALID_ARG, "job %zu: NULL data", i);
  }
  // order jobs by (log_n, dir)
  std::vector<size_t> order(count);
  for (size_t i = 0; i < count; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
    return log_ns[x] != log_ns[y] ? log_ns[x] < log_ns[y] : dirs[x] < dirs[y];
  });
  zp_ntt_batch *b = new zp_ntt_batch();
  b->count = count;
  std::vector<uint64_t *> tabs(2 * count + 1, nullptr);
  size_t scratch_need = 0;
  for (size_t i = 0; i < count; i++) {
    const size_t j = order[i];
    if (b->groups.empty() || b->groups.back().log_n != log_ns[j] || b->groups.back().dir != dirs[j])
      b->groups.push_back(NttGroup{log_ns[j], dirs[j], 0, i});
    b->groups.back().jobs++;
    if (needs_scratch(log_ns[j])) scratch_need += 1ull << log_ns[j];
  }
  int st = ZP_OK;
  if (scratch_need && hipMalloc(&b->d_scratch, scratch_need * 8) != hipSuccess) st = ZP_ERR_OUT_OF_MEMORY;
  size_t off = 0;
  for (size_t i = 0; i < count; i++) {
    const size_t j = order[i];
    tabs[i] = d_ptrs[j];
    if (needs_scratch(log_ns[j]) && b->d_scratch) {
      tabs[count + i] = b->d_scratch + off;
      off += 1ull << log_ns[j];
    }
  }
  if (st == ZP_OK && hipMalloc(&b->d_tabs, tabs.size() * sizeof(uint64_t *)) != hipSuccess) st = ZP_ERR_OUT_OF_MEMORY;
  if (st == ZP_OK &&
      hipMemcpy(b->d_tabs, tabs.data(), tabs.size() * sizeof(uint64_t *), hipMemcpyHostToDevice) != hipSuccess)
    st = ZP_ERR_HIP;
  for (const NttGroup &g : b->groups) {
    NttTables *t = nullptr;
    if (st == ZP_OK) st = make_tables(g.log_n, &t);
  }
  if (st != ZP_OK) {
    if (st == ZP_ERR_OUT_OF_MEMORY) set_last_error("zp_ntt_batch_create: device allocation failed");
    zp_ntt_batch_destroy(b);
    return st;
  }
  *out = b;
  return ZP_OK;
}

extern "C" int zp_ntt_batch_launch(zp_ntt_batch *b, int use_graph, void *stream) {
  ZP_CHECK(b != nullptr, ZP_ERR_INVALID_ARG, "zp_ntt_batch_launch: NULL batch");
  hipStream_t s = (hipStream_t)stream;
  if (!use_graph) return batch_launch_eager(b, s);
  if (!b->exec) {
    // Capture on a private stream (the null stream cannot be captured).
    hipStream_t cap = nullptr;
    ZP_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    ZP_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
    int st = batch_launch_eager(b, cap);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(cap, &g);
    (void)hipStreamDestroy(cap);
    if (st != ZP_OK) return st;
    ZP_CHECK(e == hipSuccess, ZP_ERR_HIP, "zp_ntt_batch_launch: graph capture failed: %s", hipGetErrorString(e));
    b->graph = g;
    ZP_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  }
  ZP_HIP(hipGraphLaunch(b->exec, s));
  return ZP_OK;
}

extern "C" int zp_ntt_batch_destroy(zp_ntt_batch *b) {
  if (!b) return ZP_OK;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->d_scratch) (void)hipFree(b->d_scratch);
  if (b->d_tabs) (void)hipFree(b->d_tabs);
  delete b;
  return ZP_OK;
}

extern "C" int zp_ntt_batch_device(uint64_t *const *d_ptrs, const unsigned *log_ns, const int *dirs, size_t count,
                                   void *stream) {
  zp_ntt_batch *b = nullptr;
  ZP_TRY(zp_ntt_batch_create(d_ptrs, log_ns, dirs, count, &b));
  int st = zp_ntt_batch_launch(b, 0, stream);
  if (st == ZP_OK && hipStreamSynchronize((hipStream_t)stream) != hipSuccess) {
    set_last_error("zp_ntt_batch_device: execution failed");
    st = ZP_ERR_HIP;
  }
  zp_ntt_batch_destroy(b);
  return st;
}

// ------------------------------------------------------------- utilities --
extern "C" int zp_device_malloc(void **ptr, size_t bytes) {
  ZP_CHECK(ptr != nullptr, ZP_ERR_INVALID_ARG, "zp_device_malloc: NULL out");
  if (hipMalloc(ptr, bytes ? bytes : 1) != hipSuccess) {
    set_last_error("zp_device_malloc: %zu bytes failed", bytes);
    return ZP_ERR_OUT_OF_MEMORY;
  }
  return ZP_OK;
}
extern "C" int zp_device_free(void *ptr) {
  ZP_HIP(hipFree(ptr));
  return ZP_OK;
}
extern "C" int zp_memcpy_h2d(void *dst, const void *src, size_t bytes) {
  ZP_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return ZP_OK;
}
extern "C" int zp_memcpy_d2h(void *dst, const void *src, size_t bytes) {
  ZP_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return ZP_OK;
}
extern "C" int zp_device_synchronize(void) {
  ZP_HIP(hipDeviceSynchronize());
  return ZP_OK;
}
extern "C" int zp_stream_create(void **stream) {
  ZP_CHECK(stream != nullptr, ZP_ERR_INVALID_ARG, "zp_stream_create: NULL out");
  hipStream_t s = nullptr;
  ZP_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return ZP_OK;
}
extern "C" int zp_stream_destroy(void *stream) {
  if (stream) ZP_HIP(hipStreamDestroy((hipStream_t)stream));
  return ZP_OK;
}
extern "C" int zp_stream_synchronize(void *stream) {
  ZP_HIP(hipStreamSynchronize((hipStream_t)stream));
  return ZP_OK;
}
extern "C" int zp_memcpy_h2d_async(void *dst, const void *src, size_t bytes, void *stream) {
  ZP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
  return ZP_OK;
}
extern "C" int zp_memcpy_d2h_async(void *dst, const void *src, size_t bytes, void *stream) {
  ZP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
  return ZP_OK;
}
extern "C" int zp_host_alloc(void **ptr, size_t bytes) {
  ZP_CHECK(ptr != nullptr, ZP_ERR_INVALID_ARG, "zp_host_alloc: NULL out");
  if (hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
    set_last_error("zp_host_alloc: %zu bytes failed", bytes);
    return ZP_ERR_OUT_OF_MEMORY;
  }
  return ZP_OK;
}
extern "C" int zp_host_free(void *ptr) {
  if (ptr) ZP_HIP(hipHostFree(ptr));
  return ZP_OK;
}
extern "C" int zp_device_memory_info(size_t *free_bytes, size_t *total_bytes) {
  ZP_CHECK(free_bytes && total_bytes, ZP_ERR_INVALID_ARG, "zp_device_memory_info: NULL out");
  ZP_HIP(hipMemGetInfo(free_bytes, total_bytes));
  return ZP_OK;
}
