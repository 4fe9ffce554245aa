// BLS12-377 G1 MSM: context, point preprocessing, the per-batch pipeline and
// the MSM half of the C ABI (msm_init / msm_mult of
// zprize/msm_pippenger/host/driver/driver.cpp:603-636).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <rocprim/device/device_scan.hpp>
#include <map>
#include <mutex>
#include <atomic>
#include <string>
#include <thread>
#include <vector>

#include "common.hpp"
#include "msm_host.hpp"
#include "msm_kernels.hpp"
#include "msm_plan.hpp"
#include "msm_finalize.hpp"
#include "fr253.hpp"

namespace zp {

static std::once_flag g_consts_once;
static FConsts g_K;

static const FConsts &consts() {
  std::call_once(g_consts_once, [] { g_K = make_fconsts(); });
  return g_K;
}

// ----------------------------------------------------------- kernels --
// to_r converts an input coordinate into the R' domain: K.to_int for the
// wire form x R, K.r2 for a plain x.
__global__ void k_pre_a(const g1_affine_t *in, uint64_t n, FConsts K, F to_r, F *u_out, F *v_out, F *den_out,
                        uint8_t *flag) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const g1_affine_t &p = in[i];
    if (p.infinity) {  // bls12_377_g1.h:647-651: the twisted-Edwards identity
      flag[i] = 2;
      den_out[i] = K.one;
      continue;
    }
    Packed px, py;
    const uint32_t *wx = reinterpret_cast<const uint32_t *>(p.x.data);
    const uint32_t *wy = reinterpret_cast<const uint32_t *>(p.y.data);
#pragma unroll
    for (int j = 0; j < 12; j++) {
      px.w[j] = wx[j];
      py.w[j] = wy[j];
    }
    const F x = f_mul(f_unpack(px), to_r);
    const F y = f_mul(f_unpack(py), to_r);
    // affineWeierstrassToMontgomery (bls12_377_g1.h:361-365): u = s (x - alpha), v = s y.
    const F u = f_mul(f_sub_kq(x, K.alpha, K.q2), K.mont_s);
    const F v = f_mul(K.mont_s, y);
    const F up1 = f_add(u, K.one);
    const F den = f_mul(v, up1);
    u_out[i] = u;
    v_out[i] = v;
    // affineMontgomeryToTwistedEdwards fails iff v = 0 or u = -1 (:377-388).
    if (f_is_zero_canon(den)) {
      flag[i] = 0;
      den_out[i] = K.one;
    } else {
      flag[i] = 1;
      den_out[i] = den;
    }
  }
}

// In-place inversion of a[0..n) in chunks of `chunk` per thread (Montgomery's
// trick; every element must be non-zero).  tmp holds the prefix products.
__global__ void k_batch_inv(F *a, uint64_t n, uint32_t chunk, F *tmp, FConsts K) {
  const uint64_t nchunks = (n + chunk - 1) / chunk;
  for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < nchunks;
       c += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t b0 = c * chunk, b1 = min(n, b0 + chunk);
    F acc = K.one;
    for (uint64_t i = b0; i < b1; i++) {
      acc = f_mul(acc, a[i]);
      tmp[i] = acc;
    }
    F inv = f_inv(acc, K.one);
    for (uint64_t i = b1; i-- > b0;) {
      const F prev = (i > b0) ? tmp[i - 1] : K.one;
      const F ai = a[i];
      a[i] = f_mul(inv, prev);
      inv = f_mul(inv, ai);
    }
  }
}

__global__ void k_pre_c(const F *u_in, const F *v_in, const F *inv_den, const uint8_t *flag, uint64_t n, FConsts K,
                        PackedPre *out, uint8_t *conv) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint8_t f = flag[i];
    PrePt pre;
    if (f != 1) {
      // identity (y - x, y + x, 2dxy) = (1, 1, 0); unused when not convertible
      pre.ymx = K.one;
      pre.ypx = K.one;
      pre.t2d = f_zero();
    } else {
      const F u = u_in[i], v = v_in[i], id = inv_den[i];
      const F up1 = f_add(u, K.one);
      const F inv_v = f_mul(up1, id);
      const F inv_up1 = f_mul(v, id);
      const F x = f_mul(f_mul(K.ts, u), inv_v);              // twisted_scale u / v
      const F y = f_mul(f_sub_kq(u, K.one, K.q2), inv_up1);  // (u - 1) / (u + 1)
      pre = pre_from_affine(x, y, K);
    }
    out[i] = pre_pack(pre);
    conv[i] = f != 0;
  }
}

// Window-shifted table (init option "precompute"): entry i * W + w holds
// 2^(c w) P_i in the packed precomputed form, so that every window of a
// c-bit layout adds into one shared bucket set.  Entry w is made from entry
// w - 1 by c doublings and one affine normalisation (batch inversion of Z).
// k_shift_a: entry w - 1 of points [i0, i0 + m) -> X, Y, Z after c doublings.
// A Z of zero (a point of small order whose multiple has no affine
// twisted-Edwards image) sets *bad, and the caller drops the table.
__global__ void k_shift_a(const PackedPre *tab, uint64_t i0, uint64_t m, uint32_t W, uint32_t w, uint32_t c,
                          const uint8_t *conv, FConsts K, F *X, F *Y, F *Z, int *bad) {
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < m; k += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t i = i0 + k;
    if (!conv[i]) {  // never referenced: its digits are zero
      X[k] = f_zero();
      Y[k] = K.one;
      Z[k] = K.one;
      continue;
    }
    Xyzt e = ext_from_pre(pre_unpack(tab[i * W + (w - 1)]), false, K);
    for (uint32_t d = 0; d < c; d++) ec_dbl(e, K);
    if (f_is_zero_canon(e.z)) {
      atomicOr(bad, 1);
      e.z = K.one;
    }
    X[k] = e.x;
    Y[k] = e.y;
    Z[k] = e.z;
  }
}

// k_shift_c: X, Y and 1/Z -> entry w of points [i0, i0 + m).
__global__ void k_shift_c(const F *X, const F *Y, const F *Zinv, uint64_t i0, uint64_t m, uint32_t W, uint32_t w,
                          FConsts K, PackedPre *tab) {
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < m; k += (uint64_t)gridDim.x * blockDim.x) {
    const F x = f_mul(X[k], Zinv[k]), y = f_mul(Y[k], Zinv[k]);
    tab[(i0 + k) * W + w] = pre_pack(pre_from_affine(x, y, K));
  }
}

// Entry 0 of every point: the table itself.
__global__ void k_shift_copy0(const PackedPre *pts, uint64_t n, uint32_t W, PackedPre *tab) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    tab[i * W] = pts[i];
}

// Fixed-base comb: acc = sum_j T[j][byte_j(k)] with T[j][m] = m 2^(8j) G.
__global__ void k_gen_a(const biginteger256_t *ks, uint64_t n, const PackedPre *table, FConsts K, Xyzt *acc_out,
                        F *den_out, uint8_t *flag) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    Xyzt acc = ec_identity(K);
    for (int j = 0; j < 32; j++) {
      const uint32_t m = (uint32_t)(ks[i].data[j >> 3] >> (8 * (j & 7))) & 0xff;
      if (m) ec_madd(acc, pre_unpack(table[j * 256 + m]), false, K);
    }
    acc_out[i] = acc;
    // twisted Edwards -> Montgomery needs 1 / ((Z - Y) X) (bls12_377_g1.h:389-402).
    const F den = f_mul(f_sub_kq(acc.z, acc.y, K.q2), acc.x);
    if (f_is_zero_canon(den)) {
      flag[i] = 0;
      den_out[i] = K.one;
    } else {
      flag[i] = 1;
      den_out[i] = den;
    }
  }
}

// Extended twisted Edwards + 1/((Z-Y)X) -> affine Weierstrass in the wire form.
__device__ __forceinline__ void te_to_wire(const Xyzt &a, const F &inv, const FConsts &K, biginteger384_t &xo,
                                           biginteger384_t &yo) {
  const F zpy = f_add(a.z, a.y);
  const F u = f_mul(f_mul(zpy, a.x), inv);                  // (1 + y) / (1 - y)
  const F v = f_mul(f_mul(f_mul(zpy, a.z), inv), K.ts);     // twisted_scale u / x
  const F xw = f_add(f_mul(u, K.inv_cb), K.ca_3cb);         // bls12_377_g1.h:366-374
  const F yw = f_mul(v, K.inv_cb);
  const Packed px = f_pack(f_canonical(f_mul(xw, K.to_wire)));
  const Packed py = f_pack(f_canonical(f_mul(yw, K.to_wire)));
  uint32_t *dx = reinterpret_cast<uint32_t *>(xo.data);
  uint32_t *dy = reinterpret_cast<uint32_t *>(yo.data);
#pragma unroll
  for (int j = 0; j < 12; j++) {
    dx[j] = px.w[j];
    dy[j] = py.w[j];
  }
}

__global__ void k_gen_c(const Xyzt *acc, const F *inv, const uint8_t *flag, uint64_t n, FConsts K, g1_affine_t *out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    g1_affine_t r;
    if (!flag[i]) {
#pragma unroll
      for (int j = 0; j < 6; j++) r.x.data[j] = r.y.data[j] = 0;
      r.infinity = true;
    } else {
      te_to_wire(acc[i], inv[i], K, r.x, r.y);
      r.infinity = false;
    }
    out[i] = r;
  }
}

// Diagnostics: wire-form Montgomery product through the device multiply.
__global__ void k_debug_mul(const uint64_t *a, const uint64_t *b, uint64_t *o, uint64_t n, FConsts K) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    Packed pa, pb;
    for (int j = 0; j < 6; j++) {
      pa.w[2 * j] = (uint32_t)a[6 * i + j];
      pa.w[2 * j + 1] = (uint32_t)(a[6 * i + j] >> 32);
      pb.w[2 * j] = (uint32_t)b[6 * i + j];
      pb.w[2 * j + 1] = (uint32_t)(b[6 * i + j] >> 32);
    }
    const F x = f_mul(f_unpack(pa), K.to_int), y = f_mul(f_unpack(pb), K.to_int);
    const Packed r = f_pack(f_canonical(f_mul(f_mul(x, y), K.to_wire)));
    for (int j = 0; j < 6; j++) o[6 * i + j] = (uint64_t)r.w[2 * j] | ((uint64_t)r.w[2 * j + 1] << 32);
  }
}

__global__ void k_debug_pre_wire(const PackedPre *pts, uint64_t i0, uint64_t n, FConsts K, uint64_t *o) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const PrePt p = pre_unpack(pts[i0 + i]);
    const F v[3] = {p.ymx, p.ypx, p.t2d};
    for (int c = 0; c < 3; c++) {
      const Packed r = f_pack(f_canonical(f_mul(v[c], K.to_wire)));
      for (int j = 0; j < 6; j++) o[18 * i + 6 * c + j] = (uint64_t)r.w[2 * j] | ((uint64_t)r.w[2 * j + 1] << 32);
    }
  }
}

// Fill kernel (used instead of hipMemsetAsync so the sequence can be graph-captured).
__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

static unsigned grid_for(uint64_t n, unsigned threads = MSM_THREADS, unsigned cap = 65535) {
  uint64_t g = (n + threads - 1) / threads;
  if (g == 0) g = 1;
  return (unsigned)std::min<uint64_t>(g, cap);
}

// --------------------------------------------------------- workspace --
template <typename T>
static int grow(T *&p, size_t &cap, size_t need) {
  if (need <= cap && p) return ZP_OK;
  if (p) (void)hipFree(p);
  p = nullptr;
  cap = 0;
  if (hipMalloc(&p, std::max<size_t>(need, 1) * sizeof(T)) != hipSuccess) {
    p = nullptr;
    set_last_error("MSM workspace: allocation of %zu bytes failed", need * sizeof(T));
    return ZP_ERR_OUT_OF_MEMORY;
  }
  cap = need;
  return ZP_OK;
}

}  // namespace zp

using namespace zp;

// Per-batch device working memory; one workspace serves one batch at a time.
struct zp_msm_workspace {
  size_t cap_K = 0, cap_K2 = 0, cap_K3 = 0, cap_K4 = 0, cap_K5 = 0;
  size_t cap_entries = 0, cap_units = 0, cap_units2 = 0, cap_groups = 0, cap_gs = 0, cap_win = 0, cap_wsum = 0;
  size_t cap_jobs = 0, cap_pt0 = 0, cap_b0 = 0, cap_scan = 0;
  size_t cap_digits = 0, cap_counts = 0, cap_scanned = 0, cap_segs = 0, cap_tmp = 0, cap_wg0 = 0;
  int32_t *digits = nullptr;
  uint32_t *tmp = nullptr;      // two-level sort: pass A output
  uint32_t *win_g0 = nullptr;   // two-level sort: first group of each window
  uint32_t *counts = nullptr, *scanned = nullptr;
  SegDev *segs = nullptr;
  uint32_t *cnt = nullptr, *off = nullptr, *cursor = nullptr, *nunits = nullptr, *ustart = nullptr;
  uint32_t *ubucket = nullptr, *vals = nullptr, *gstart = nullptr, *pt0 = nullptr, *b0 = nullptr;
  Xyzt *partial = nullptr, *groups = nullptr, *winsum = nullptr;
  Xyzt *winpart = nullptr;  // slice sums of windows with many groups
  size_t cap_winpart = 0;
  uint32_t *umax = nullptr;  // largest unit count of any bucket (k_units_per_bucket)
  size_t cap_umax = 0;
  MsmJobDev *jobs = nullptr;
  WinDev *wins = nullptr;
  uint8_t *scan_tmp = nullptr;
  std::vector<Xyzt> h_winsum;
  uint64_t *nc_sums = nullptr;  // 5 words per exceptional-point group per job
  size_t cap_nc = 0;
  std::vector<uint64_t> h_nc_sums;
  FinJobDev *fin = nullptr;  // per-job finalisation descriptors
  size_t cap_fin = 0;
  uint64_t *d_out = nullptr;  // per-job results, 18 words (g1_projective_t) each
  size_t cap_out = 0;
  std::vector<uint64_t> h_out;
  FConsts *d_consts = nullptr;  // the field constants in device memory (k_msm_finalize)
  size_t cap_consts = 0;
  // Arena mode (zp_msm_mult_prefix_noalloc): every buffer is carved from
  // [arena, arena + arena_cap) instead of allocated; release() frees nothing.
  char *arena = nullptr;
  size_t arena_cap = 0, arena_used = 0;

  void release() {
    if (arena) {
      char *a = arena;
      const size_t c = arena_cap;
      *this = zp_msm_workspace();
      arena = a;
      arena_cap = c;
      return;
    }
    void *ps[] = {cnt, off, cursor, nunits, ustart, ubucket, vals, gstart, pt0, b0, partial, groups,
                  winsum, jobs, wins, scan_tmp, digits, counts, scanned, segs, nc_sums, tmp, win_g0,
                  winpart, umax, fin, d_out, d_consts};
    for (void *p : ps)
      if (p) (void)hipFree(p);
    *this = zp_msm_workspace();
  }
};

// Table points without a twisted-Edwards image (driver.cpp:160-172), grouped
// by value: the curve has only five such points, however many times a table
// repeats them.  Their digits are skipped by the bucket pipeline (their
// precomputed entry is the identity); per batch k_nc_sums adds up each
// group's scalars exactly on the GPU and the host adds S_g * P_g once per
// group (driver.cpp:349-372 does one double-and-add per point).
struct NcGroup {
  F x, y;                     // Weierstrass affine, R' Montgomery form, canonical
  std::vector<uint32_t> idx;  // table indices, ascending
};

// Host (pageable) scalars -> device through pinned staging.  The caller's
// array is split into kLanes contiguous slices; each slice is copied by its
// own host thread into one of that lane's two pinned slabs, and the slab is
// sent by DMA on the lane's stream while the thread fills the other slab.
// A plain pageable hipMemcpy runs at ~25 GB/s here; this keeps the PCIe link
// busy from several DMA queues.  upload() returns once the data is on the
// device.  One upload at a time per uploader.
struct HostUploader {
  static constexpr int kLanes = 8;
  static constexpr size_t kSlab = size_t(16) << 20;
  struct Lane {
    void *slab[2] = {nullptr, nullptr};
    hipEvent_t sent[2] = {nullptr, nullptr};
    hipStream_t s = nullptr;
  };
  Lane lanes[kLanes];
  bool ready = false;

  int init() {
    if (ready) return ZP_OK;
    for (Lane &l : lanes) {
      for (int k = 0; k < 2; k++) {
        ZP_HIP(hipHostMalloc(&l.slab[k], kSlab, hipHostMallocDefault));
        ZP_HIP(hipEventCreateWithFlags(&l.sent[k], hipEventDisableTiming));
      }
      ZP_HIP(hipStreamCreateWithFlags(&l.s, hipStreamNonBlocking));
    }
    ready = true;
    return ZP_OK;
  }

  void release() {
    for (Lane &l : lanes) {
      if (l.s) (void)hipStreamSynchronize(l.s);
      for (int k = 0; k < 2; k++) {
        if (l.slab[k]) (void)hipHostFree(l.slab[k]);
        if (l.sent[k]) (void)hipEventDestroy(l.sent[k]);
      }
      if (l.s) (void)hipStreamDestroy(l.s);
      l = Lane();
    }
    ready = false;
  }

  int upload(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return ZP_OK;
    if (bytes < 4 * kSlab) {  // small: one pageable copy
      ZP_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
      return ZP_OK;
    }
    ZP_TRY(init());
    int dev = 0;
    ZP_HIP(hipGetDevice(&dev));
    const size_t per = ((bytes + kLanes - 1) / kLanes + 4095) & ~size_t(4095);
    std::atomic<int> failed{0};
    std::vector<std::thread> th;
    for (int t = 0; t < kLanes; t++) {
      const size_t b0 = std::min(bytes, per * t), b1 = std::min(bytes, b0 + per);
      if (b0 >= b1) break;
      th.emplace_back([&, t, b0, b1] {
        Lane &l = lanes[t];
        bool ok = hipSetDevice(dev) == hipSuccess;
        bool used[2] = {false, false};
        int k = 0;
        for (size_t o = b0; ok && o < b1; o += kSlab, k ^= 1) {
          const size_t len = std::min(kSlab, b1 - o);
          if (used[k]) ok = hipEventSynchronize(l.sent[k]) == hipSuccess;  // slab k's last DMA done
          if (!ok) break;
          memcpy(l.slab[k], (const char *)src + o, len);
          ok = hipMemcpyAsync((char *)dst + o, l.slab[k], len, hipMemcpyHostToDevice, l.s) == hipSuccess &&
               hipEventRecord(l.sent[k], l.s) == hipSuccess;
          used[k] = true;
        }
        if (hipStreamSynchronize(l.s) != hipSuccess) ok = false;
        if (!ok) failed = 1;
      });
    }
    for (auto &x : th) x.join();
    ZP_CHECK(!failed, ZP_ERR_HIP, "scalar upload through pinned staging failed");
    return ZP_OK;
  }

  // Device -> pageable host, the reverse pipeline: each lane DMAs its slice
  // slab by slab into its pinned slabs (after `after` completes on its
  // stream) while its thread copies the previous slab out.
  int download(void *dst, const void *src, size_t bytes, hipStream_t after) {
    if (bytes == 0) return ZP_OK;
    if (after) ZP_HIP(hipStreamSynchronize(after));
    if (bytes < 4 * kSlab) {
      ZP_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
      return ZP_OK;
    }
    ZP_TRY(init());
    int dev = 0;
    ZP_HIP(hipGetDevice(&dev));
    const size_t per = ((bytes + kLanes - 1) / kLanes + 4095) & ~size_t(4095);
    std::atomic<int> failed{0};
    std::vector<std::thread> th;
    for (int t = 0; t < kLanes; t++) {
      const size_t b0 = std::min(bytes, per * t), b1 = std::min(bytes, b0 + per);
      if (b0 >= b1) break;
      th.emplace_back([&, t, b0, b1] {
        Lane &l = lanes[t];
        bool ok = hipSetDevice(dev) == hipSuccess;
        auto issue = [&](size_t o, int k) {
          const size_t len = std::min(kSlab, b1 - o);
          return hipMemcpyAsync(l.slab[k], (const char *)src + o, len, hipMemcpyDeviceToHost, l.s) == hipSuccess &&
                 hipEventRecord(l.sent[k], l.s) == hipSuccess;
        };
        int k = 0;
        if (ok) ok = issue(b0, 0);
        for (size_t o = b0; ok && o < b1; o += kSlab, k ^= 1) {
          if (o + kSlab < b1) ok = issue(o + kSlab, k ^ 1);  // next slab in flight while this one is copied
          if (ok) ok = hipEventSynchronize(l.sent[k]) == hipSuccess;
          if (ok) memcpy((char *)dst + o, l.slab[k], std::min(kSlab, b1 - o));
        }
        if (hipStreamSynchronize(l.s) != hipSuccess) ok = false;
        if (!ok) failed = 1;
      });
    }
    for (auto &x : th) x.join();
    ZP_CHECK(!failed, ZP_ERR_HIP, "download through pinned staging failed");
    return ZP_OK;
  }
};

// Stand-alone host <-> device staging for callers with large pageable
// arrays (the serving loop's job inputs and NTT results).
struct zp_transfer {
  HostUploader u;
};

struct zp_msm_context {
  size_t n = 0;
  MsmOptions opts;
  hipStream_t stream = nullptr;
  PackedPre *d_pts = nullptr;
  uint8_t *d_conv = nullptr;
  std::vector<NcGroup> nc_groups;
  size_t num_nonconv = 0;
  uint32_t *d_nc_idx = nullptr;  // every group's indices, group after group
  uint32_t *d_nc_off = nullptr;  // nc_groups.size() + 1 offsets into d_nc_idx
  NcMultDev *d_nc_mult = nullptr;  // P, 2P, 3P of each group (k_msm_finalize)
  // Window-shifted table (init option "precompute"): d_shift[i * shift_W + w]
  // = 2^(shift_c w) P_i; null when not requested.
  PackedPre *d_shift = nullptr;
  int shift_c = 0, shift_W = 0;
  zp_msm_workspace ws;
  uint64_t *d_scalars = nullptr;  // staging for host scalars
  size_t cap_scalars = 0;
  uint64_t *d_scalars2 = nullptr;  // second buffer: the next batch uploads while one computes
  size_t cap_scalars2 = 0;
  HostUploader uploader;  // host scalars of msm_mult
};

namespace zp {

// A workspace buffer of at least `need` elements: allocated (kept and reused
// while large enough), or in arena mode carved from the arena, 256-byte
// aligned, with ZP_ERR_TOO_LARGE when the arena is exhausted.
template <typename T>
static int wgrow(zp_msm_workspace &ws, T *&p, size_t &cap, size_t need) {
  if (!ws.arena) return grow(p, cap, need);
  const size_t bytes = (std::max<size_t>(need, 1) * sizeof(T) + 255) & ~size_t(255);
  ZP_CHECK(ws.arena_used + bytes <= ws.arena_cap, ZP_ERR_TOO_LARGE,
           "MSM scratch: the job needs more than the %zu-byte scratch array (%zu bytes used, %zu more asked)",
           ws.arena_cap, ws.arena_used, bytes);
  p = reinterpret_cast<T *>(ws.arena + ws.arena_used);
  ws.arena_used += bytes;
  cap = need;
  return ZP_OK;
}

static uint32_t auto_log_u(size_t entries, int requested) {
  if (requested >= 0) return (uint32_t)requested;
  // aim for ~2^19 accumulation slices of 8 .. 512 entries (longer slices
  // cut fewer buckets into several units for k_unit_tree to add; at 2^10
  // entries the 320-job mixed batch was ~2 % slower, the 2^26 MSM equal)
  int lu = 0;
  while ((entries >> (lu + 1)) >= (1u << 19) && lu < 9) lu++;
  return (uint32_t)std::max(lu, 3);
}

static uint32_t auto_log_g(uint64_t K, int requested) {
  if (requested >= 0) return (uint32_t)requested;
  // ~16K groups in total, 4 .. 64 buckets per group
  uint32_t lg = 2;
  while (lg < 6 && (K >> lg) > 16384) lg++;
  return lg;
}

// Exact 320-bit sums of the scalars of each exceptional-point group (one
// block per group), over the group's table indices below the job's prefix n.
__global__ void __launch_bounds__(256) k_nc_sums(const uint64_t *scalars, const uint32_t *idx,
                                                 const uint32_t *goff, uint64_t n, uint64_t *out) {
  __shared__ uint64_t red[256][5];
  const uint32_t g = blockIdx.x, tid = threadIdx.x;
  uint64_t acc[5] = {0, 0, 0, 0, 0};
  for (uint32_t e = goff[g] + tid; e < goff[g + 1]; e += 256) {
    const uint32_t i = idx[e];
    if (i >= n) continue;
    uint64_t c = 0;
#pragma unroll
    for (int w = 0; w < 4; w++) {
      const uint64_t a = acc[w] + c;
      const uint64_t c1 = a < c;
      acc[w] = a + scalars[4 * (uint64_t)i + w];
      c = c1 + (acc[w] < a);
    }
    acc[4] += c;
  }
#pragma unroll
  for (int w = 0; w < 5; w++) red[tid][w] = acc[w];
  __syncthreads();
  for (uint32_t h = 128; h > 0; h >>= 1) {
    if (tid < h) {
      uint64_t c = 0;
#pragma unroll
      for (int w = 0; w < 5; w++) {
        const uint64_t a = red[tid][w] + c;
        const uint64_t c1 = a < c;
        red[tid][w] = a + red[tid + h][w];
        c = c1 + (red[tid][w] < a);
      }
    }
    __syncthreads();
  }
  if (tid < 5) out[5 * (uint64_t)g + tid] = red[0][tid];
}

// Window combine (driver.cpp:340-347 as Horner), conversion to affine
// Weierstrass (bls12_377_g1.h:427-439) and the non-convertible fallback
// (driver.cpp:349-372), then the wire form (x R, y R, R) or z = 0.
static void finalize(const HostLayout &L, const Xyzt *win, const std::vector<NcGroup> &groups,
                     const uint64_t *nc_sums, g1_projective_t *out) {
  const FConsts &K = consts();
  Xyzt acc = win[L.W - 1];
  for (int w = L.W - 2; w >= 0; w--) {
    for (int i = 0; i < L.width[w]; i++) ec_dbl(acc, K);
    ec_add(acc, win[w], K);
  }
  Jac res;
  res.Z = f_zero();
  res.X = f_zero();
  res.Y = K.one;
  const F den = f_mul(f_sub_kq(acc.z, acc.y, K.q2), acc.x);
  if (!f_is_zero_canon(acc.z) && !f_is_zero_canon(den)) {
    const F inv = f_inv(den, K.one);
    const F zpy = f_add(acc.z, acc.y);
    const F u = f_mul(f_mul(zpy, acc.x), inv);
    const F v = f_mul(f_mul(f_mul(zpy, acc.z), inv), K.ts);
    res.X = f_canonical(f_add(f_mul(u, K.inv_cb), K.ca_3cb));
    res.Y = f_canonical(f_mul(v, K.inv_cb));
    res.Z = K.one;
  }
  // S_g * P_g for each group of exceptional points: S_g is the plain integer
  // sum of the group's scalars (no reduction mod r: these points lie outside
  // the order-r subgroup), up to 2^288.
  for (size_t g = 0; g < groups.size(); g++) {
    const uint64_t *S = nc_sums + 5 * g;
    int top = -1;
    for (int bit = 319; bit >= 0; bit--)
      if ((S[bit >> 6] >> (bit & 63)) & 1) {
        top = bit;
        break;
      }
    if (top < 0) continue;
    Jac kp;
    kp.X = f_zero();
    kp.Y = K.one;
    kp.Z = f_zero();
    for (int bit = top; bit >= 0; bit--) {
      jac_dbl(kp, K);
      if ((S[bit >> 6] >> (bit & 63)) & 1) jac_add_aff(kp, groups[g].x, groups[g].y, K);
    }
    jac_add_jac(res, kp, K);
  }
  memset(out, 0, sizeof(*out));
  if (f_is_zero_canon(res.Z)) return;  // infinity: z = 0
  const F zi = f_inv(res.Z, K.one), zi2 = f_sqr(zi);
  const F x = f_mul(res.X, zi2), y = f_mul(res.Y, f_mul(zi2, zi));
  const U384 wx = u384_from_f(f_mul(x, K.to_wire));
  const U384 wy = u384_from_f(f_mul(y, K.to_wire));
  const U384 wz = u384_from_f(f_mul(K.one, K.to_wire));
  memcpy(out->x.data, wx.w, 48);
  memcpy(out->y.data, wy.w, 48);
  memcpy(out->z.data, wz.w, 48);
}

// Whether a job of n points over table cx takes the window-shifted table
// (ONE window of n * W virtual points, every window adding into one shared
// bucket set of 2^(c-1) buckets): only when the table has the copies, no
// window is forced, and it costs fewer bucket additions than the per-window
// layout `per_window` (small prefixes of a big table do not: the shared
// bucket set outweighs their points).  LS receives the shifted layout.
static bool job_uses_shifted(const zp_msm_context *cx, size_t n, bool window_forced, const HostLayout &per_window,
                             HostLayout &LS) {
  if (!cx->d_shift || window_forced || !n) return false;
  LS = layout_uniform(cx->shift_c);
  layout_finish(LS);
  return msm_cost(n, LS, true) < msm_cost(n, per_window, false);
}

// Host description of one job of a batch.
struct HostJob {
  zp_msm_context *ctx;
  size_t n;
  const uint64_t *d_scalars;
  g1_projective_t *out;
  int window = 0;  // per-job window override (autotuner), 0 = none
};

// Everything the launch sequence of one batch needs, decided on the host.
struct BatchPlan {
  std::vector<HostJob> jobs;
  std::vector<HostLayout> layouts;
  std::vector<MsmJobDev> jd;
  std::vector<WinDev> wins;
  std::vector<uint32_t> pt0, b0, gstart, win_g0;
  std::vector<SegDev> segs;
  std::vector<size_t> nc0;
  std::vector<FinJobDev> fin;
  uint64_t K = 0, entries = 0, points = 0, counters = 0, n_max = 0, GT = 0, max_units = 0;
  uint32_t max_nb = 0, max_lbits = 0, max_groups = 1, max_chunks = 1, Wt = 0, log_u = 0, log_g = 0, gs = 0, S = 1;
  bool any_pre = false, sorted_path = false, msd_path = false, host_fin = false;
  BatchDev B;
};

// Plans a batch: layouts, windows, sort groups and work-unit sizes; grows
// the workspace and uploads the batch's tables on `stream`.  Nothing here is
// recorded by a graph capture of launch_batch.
static int plan_batch(zp_msm_workspace &ws, const HostJob *jobs, size_t count, const MsmOptions &bopts,
                      hipStream_t stream, BatchPlan &P) {
  P.jobs.assign(jobs, jobs + count);
  P.layouts.assign(count, HostLayout());
  P.jd.assign(count, MsmJobDev());
  P.pt0.assign(count, 0);
  P.b0.assign(count, 0);
  std::vector<HostLayout> &layouts = P.layouts;
  std::vector<MsmJobDev> &jd = P.jd;
  std::vector<WinDev> &wins = P.wins;
  std::vector<uint32_t> &pt0 = P.pt0, &b0 = P.b0, &gstart = P.gstart;
  std::vector<SegDev> &segs = P.segs;
  uint64_t &K = P.K, &entries = P.entries, &points = P.points, &counters = P.counters, &n_max = P.n_max;
  uint32_t &max_nb = P.max_nb;
  bool &any_pre = P.any_pre;
  for (size_t j = 0; j < count; j++) {
    const HostJob &h = jobs[j];
    MsmOptions o = h.ctx->opts;
    if (bopts.window || bopts.reference) {
      o.window = bopts.window;
      o.reference = bopts.reference;
    }
    if (h.window) {
      o.window = h.window;
      o.reference = false;
    }
    layouts[j] = choose_layout(o, h.n);
    const zp_msm_context *cx = h.ctx;
    // A table with window-shifted copies serves the job as ONE window of
    // n * W virtual points (digit w of point i = virtual point i * W + w, the
    // table entry 2^(c w) P_i) when that costs fewer bucket additions than the
    // per-window layout (small prefixes of a big table do not: one shared
    // bucket set of 2^(c-1) buckets outweighs their points).
    HostLayout LS;
    const bool pre = job_uses_shifted(cx, h.n, h.window || bopts.window || bopts.reference, layouts[j], LS);
    MsmJobDev &d = jd[j];
    d.conv = cx->d_conv;
    d.scalars = h.d_scalars;
    d.pt0 = (uint32_t)points;
    d.win0 = (uint32_t)wins.size();
    d.bucket0 = (uint32_t)K;
    d.dn = (uint32_t)h.n;
    d.pre = pre;
    pt0[j] = d.pt0;
    b0[j] = d.bucket0;
    if (pre) {
      any_pre = true;
      uint32_t nbv = 0;
      for (int w = 0; w < LS.W; w++) nbv = std::max(nbv, LS.nb[w]);
      WinDev v{};
      v.job = (uint32_t)j;
      v.width = (uint32_t)cx->shift_c;
      v.nb = nbv;
      v.base = (uint32_t)K;
      wins.push_back(v);
      K += nbv;
      max_nb = std::max(max_nb, nbv);
      d.W = 1;
      // the layout's windows, for the digits only (no buckets)
      d.dwin0 = (uint32_t)wins.size();
      d.dW = (uint32_t)LS.W;
      for (int w = 0; w < LS.W; w++) {
        WinDev wd{};
        wd.job = (uint32_t)j;
        wd.off = (uint32_t)LS.off[w];
        wd.width = (uint32_t)LS.width[w];
        wd.base = (uint32_t)K;
        wd.top = w == LS.W - 1;
        wins.push_back(wd);
      }
      d.pts = cx->d_shift;
      d.n = (uint32_t)(h.n * LS.W);
      HostLayout L1;  // the result is the one window's sum
      L1.W = 1;
      L1.width[0] = cx->shift_c;
      L1.nb[0] = nbv;
      layouts[j] = L1;
      // One window of up to 2^23 buckets: the two-level sort splits it into
      // up to 2^14 groups (see below) and takes 2^18-entry chunks, so that
      // pass A still writes runs of ~16 entries per group and chunk, and
      // pass B's groups stay small enough for L2 to merge its scattered writes.
      d.chunks = (uint32_t)std::min<uint64_t>(16384, std::max<uint64_t>(1, (uint64_t)d.n >> 18));
    } else {
      const HostLayout &L = layouts[j];
      d.pts = cx->d_pts;
      d.n = (uint32_t)h.n;
      d.W = (uint32_t)L.W;
      d.dwin0 = d.win0;
      d.dW = d.W;
      for (int w = 0; w < L.W; w++) {
        WinDev wd{};
        wd.job = (uint32_t)j;
        wd.off = (uint32_t)L.off[w];
        wd.width = (uint32_t)L.width[w];
        wd.nb = L.nb[w];
        wd.base = (uint32_t)K;
        wd.top = w == L.W - 1;
        K += L.nb[w];
        max_nb = std::max(max_nb, L.nb[w]);
        wins.push_back(wd);
      }
      // sorted path: C chunks of <= chunk_len points per window.  Chunks of
      // 2^16 points (up to 256 of them) keep each scatter block's output small
      // enough for L2 to merge its partial-line writes (measured at 2^26:
      // 64 chunks 179 ms, 128 174, 256 170-175, 512 174).
      // (beyond 2^26 points, chunks of 2^18: pass A stores a chunk offset in
      // 31 - lbits bits)
      d.chunks = (uint32_t)std::max<uint64_t>(1, std::max<uint64_t>(std::min<uint64_t>(256, h.n >> 16), h.n >> 18));
    }
    d.chunk_len = (uint32_t)std::max<uint64_t>(1, ((uint64_t)d.n + d.chunks - 1) / d.chunks);
    d.dig0 = entries;
    for (uint32_t w = 0; w < d.W; w++)
      for (uint32_t c = 0; c < d.chunks; c++) segs.push_back(SegDev{(uint32_t)j, d.win0 + w, c});
    points += h.n;
    entries += (uint64_t)d.n * d.W;
    n_max = std::max<uint64_t>(n_max, d.n);
  }
  // Sort: the single-level LDS histogram holds one window's buckets (<= 2^15
  // u32 = 128 KiB); the two-level sort takes any width (pass A: <= 2^12
  // groups per window in LDS, pass B: 2^lbits buckets per group).
  const char *force_atomic = getenv("ZP_MSM_ATOMIC_SORT");
  int sort = bopts.sort ? bopts.sort : jobs[0].ctx->opts.sort;
  if (force_atomic && atoi(force_atomic)) sort = 3;
  if (sort == 0) sort = max_nb <= 32768 ? 1 : 2;
  if (sort == 1 && max_nb > 32768) sort = 2;
  if (any_pre) sort = 2;  // a one-window job over shifted tables needs the two-level sort
  P.sorted_path = sort == 1;
  P.msd_path = sort == 2;
  const bool sorted_path = P.sorted_path, msd_path = P.msd_path;
  // Two-level split per window: group = bucket >> lbits with about 2^11
  // groups per window (measured at 2^26, window 20: 2^11 groups 157 ms, 2^9
  // 190 ms, 2^7 302 ms: pass B is one block per group), and at most 2^12
  // (pass A's LDS histogram); a few-bucket top window gets one bucket per group.
  const int lb_opt = bopts.sort_lbits >= 0 ? bopts.sort_lbits : jobs[0].ctx->opts.sort_lbits;
  uint64_t &GT = P.GT;
  uint32_t &max_lbits = P.max_lbits, &max_groups = P.max_groups, &max_chunks = P.max_chunks;
  for (size_t j = 0; j < count; j++) max_chunks = std::max(max_chunks, jd[j].chunks);
  for (size_t j = 0; j < count; j++) {
    MsmJobDev &d = jd[j];
    d.cnt0 = counters;
    d.g0 = (uint32_t)GT;
    uint64_t kj = 0;
    for (uint32_t w = d.win0; w < d.win0 + d.W; w++) {
      WinDev &wd = wins[w];
      uint32_t log_nb = 0;
      while ((1u << log_nb) < wd.nb) log_nb++;
      // 2^11 groups per window (a shifted-table job's one big window too):
      // pass A's tiles of ~2^15 entries then write runs of ~16 entries per
      // group, and pass B places a group of any size tile by tile.
      const uint32_t hb = d.pre ? 11 : 12, hd = 11;
      const uint32_t lo = log_nb > hb ? log_nb - hb : 0;
      wd.lbits = std::max<uint32_t>(lo, lb_opt >= 0 ? (uint32_t)lb_opt : (log_nb > hd ? log_nb - hd : 0));
      wd.g0 = (uint32_t)GT;
      ZP_CHECK(wd.lbits == 0 || ((uint64_t)d.chunk_len << (wd.lbits + 1)) <= (1ull << 32), ZP_ERR_TOO_LARGE,
               "MSM batch: chunk of %u entries with %u low bucket bits", d.chunk_len, wd.lbits);
      const uint32_t H = bucket_groups(wd.nb, wd.lbits);
      GT += H;
      max_groups = std::max(max_groups, H);
      max_lbits = std::max(max_lbits, wd.lbits);
      kj += wd.nb;
    }
    // digit-only windows of a shifted-table job: no groups (group lookups
    // take the last window whose first group is <= g, the next sort window)
    for (uint32_t w = d.dwin0; d.pre && w < d.dwin0 + d.dW; w++) {
      wins[w].g0 = (uint32_t)GT;
      wins[w].lbits = 0;
    }
    counters += (msd_path ? GT - d.g0 : kj) * d.chunks;
  }
  ZP_CHECK(points < (1ull << 31) && K < (1ull << 31) && entries < (1ull << 32), ZP_ERR_TOO_LARGE,
           "MSM batch too large (%llu points, %llu buckets)", (unsigned long long)points, (unsigned long long)K);
  const uint32_t Wt = P.Wt = (uint32_t)wins.size();
  const uint32_t log_u = P.log_u = auto_log_u(entries, bopts.log_u >= 0 ? bopts.log_u : jobs[0].ctx->opts.log_u);
  const uint32_t log_g = P.log_g = auto_log_g(K, bopts.log_g >= 0 ? bopts.log_g : jobs[0].ctx->opts.log_g);
  uint32_t &gs = P.gs;
  gs = 0;
  for (uint32_t w = 0; w < Wt; w++) {
    gstart.push_back(gs);
    gs += (wins[w].nb + (1u << log_g) - 1) >> log_g;
  }
  gstart.push_back(gs);
  const uint64_t max_units = P.max_units = (entries >> log_u) + K + 1;
  ZP_CHECK(max_units < (1ull << 32), ZP_ERR_TOO_LARGE, "MSM batch: too many work units");
  // ---- workspace
  ZP_TRY(wgrow(ws, ws.cnt, ws.cap_K, K + 1));
  ZP_TRY(wgrow(ws, ws.off, ws.cap_K2, K + 1));
  ZP_TRY(wgrow(ws, ws.cursor, ws.cap_K3, K + 1));
  ZP_TRY(wgrow(ws, ws.nunits, ws.cap_K4, K + 1));
  ZP_TRY(wgrow(ws, ws.ustart, ws.cap_K5, K + 1));
  ZP_TRY(wgrow(ws, ws.vals, ws.cap_entries, std::max<uint64_t>(entries, 1)));
  ZP_TRY(wgrow(ws, ws.ubucket, ws.cap_units, max_units));
  ZP_TRY(wgrow(ws, ws.partial, ws.cap_units2, max_units));
  ZP_TRY(wgrow(ws, ws.groups, ws.cap_groups, (size_t)gs + 1));
  ZP_TRY(wgrow(ws, ws.gstart, ws.cap_gs, gstart.size()));
  ZP_TRY(wgrow(ws, ws.wins, ws.cap_win, wins.size()));
  ZP_TRY(wgrow(ws, ws.winsum, ws.cap_wsum, wins.size()));
  ZP_TRY(wgrow(ws, ws.jobs, ws.cap_jobs, count));
  ZP_TRY(wgrow(ws, ws.pt0, ws.cap_pt0, count));
  ZP_TRY(wgrow(ws, ws.b0, ws.cap_b0, count));
  size_t scan_bytes = 0, scan_bytes2 = 0;
  ZP_HIP(rocprim::exclusive_scan(nullptr, scan_bytes, ws.cnt, ws.off, 0u, (size_t)K + 1, rocprim::plus<uint32_t>(),
                                 stream));
  if (msd_path) {
    ZP_TRY(wgrow(ws, ws.tmp, ws.cap_tmp, std::max<uint64_t>(entries, 1)));
    ZP_TRY(wgrow(ws, ws.win_g0, ws.cap_wg0, std::max<size_t>(wins.size(), 1)));
  }
  if (sorted_path || msd_path) {
    ZP_TRY(wgrow(ws, ws.digits, ws.cap_digits, std::max<uint64_t>(entries, 1)));
    ZP_TRY(wgrow(ws, ws.counts, ws.cap_counts, counters + 1));
    ZP_TRY(wgrow(ws, ws.scanned, ws.cap_scanned, counters + 1));
    ZP_TRY(wgrow(ws, ws.segs, ws.cap_segs, std::max<size_t>(segs.size(), 1)));
    ZP_HIP(rocprim::exclusive_scan(nullptr, scan_bytes2, ws.counts, ws.scanned, 0u, (size_t)counters + 1,
                                   rocprim::plus<uint32_t>(), stream));
  }
  ZP_TRY(wgrow(ws, ws.scan_tmp, ws.cap_scan, std::max<size_t>(std::max(scan_bytes, scan_bytes2), 16)));
  ZP_HIP(hipMemcpyAsync(ws.jobs, jd.data(), count * sizeof(MsmJobDev), hipMemcpyHostToDevice, stream));
  ZP_HIP(hipMemcpyAsync(ws.wins, wins.data(), wins.size() * sizeof(WinDev), hipMemcpyHostToDevice, stream));
  ZP_HIP(hipMemcpyAsync(ws.pt0, pt0.data(), count * 4, hipMemcpyHostToDevice, stream));
  ZP_HIP(hipMemcpyAsync(ws.b0, b0.data(), count * 4, hipMemcpyHostToDevice, stream));
  ZP_HIP(hipMemcpyAsync(ws.gstart, gstart.data(), gstart.size() * 4, hipMemcpyHostToDevice, stream));
  if ((sorted_path || msd_path) && !segs.empty())
    ZP_HIP(hipMemcpyAsync(ws.segs, segs.data(), segs.size() * sizeof(SegDev), hipMemcpyHostToDevice, stream));
  P.win_g0.resize(wins.size());
  for (size_t w = 0; w < wins.size(); w++) P.win_g0[w] = wins[w].g0;
  if (msd_path && !wins.empty())
    ZP_HIP(hipMemcpyAsync(ws.win_g0, P.win_g0.data(), wins.size() * 4, hipMemcpyHostToDevice, stream));
  BatchDev &B = P.B;
  B.jobs = ws.jobs;
  B.job_pt0 = ws.pt0;
  B.job_b0 = ws.b0;
  B.wins = ws.wins;
  B.J = (uint32_t)count;
  B.total_points = (uint32_t)points;
  B.K = (uint32_t)K;
  // reduction slices: windows of many groups (a shifted-table job's one window)
  uint32_t max_wgroups = 0;
  for (uint32_t w = 0; w < Wt; w++) max_wgroups = std::max(max_wgroups, gstart[w + 1] - gstart[w]);
  P.S = max_wgroups > 8192 ? 32 : 1;
  if (P.S > 1) ZP_TRY(wgrow(ws, ws.winpart, ws.cap_winpart, (size_t)Wt * P.S));
  ZP_TRY(wgrow(ws, ws.umax, ws.cap_umax, 1));
  // exceptional-point groups, per job
  P.nc0.assign(count + 1, 0);
  for (size_t j = 0; j < count; j++) P.nc0[j + 1] = P.nc0[j] + jobs[j].ctx->nc_groups.size();
  if (P.nc0[count]) ZP_TRY(wgrow(ws, ws.nc_sums, ws.cap_nc, 5 * P.nc0[count]));
  P.host_fin = bopts.host_finalize || jobs[0].ctx->opts.host_finalize;
  // device finalisation: per-job descriptors, constants, results
  P.fin.resize(count);
  for (size_t j = 0; j < count; j++) {
    P.fin[j].win0 = jd[j].win0;
    P.fin[j].W = jd[j].W;
    P.fin[j].nc0 = (uint32_t)P.nc0[j];
    P.fin[j].ng = (uint32_t)jobs[j].ctx->nc_groups.size();
    P.fin[j].mult = jobs[j].ctx->d_nc_mult;
  }
  if (!ws.d_consts) {
    ZP_TRY(wgrow(ws, ws.d_consts, ws.cap_consts, 1));
    ZP_HIP(hipMemcpy(ws.d_consts, &consts(), sizeof(FConsts), hipMemcpyHostToDevice));
  }
  ZP_TRY(wgrow(ws, ws.fin, ws.cap_fin, count));
  ZP_TRY(wgrow(ws, ws.d_out, ws.cap_out, 18 * count));
  ZP_HIP(hipMemcpyAsync(ws.fin, P.fin.data(), count * sizeof(FinJobDev), hipMemcpyHostToDevice, stream));
  return ZP_OK;
}

// The launch sequence of a planned batch: kernels and device-to-device work
// only (no allocation, no host synchronisation, no memset), so a HIP graph
// can capture it.  With device finalisation the results are copied to
// h_out (18 words per job; pinned memory when captured).
static int launch_batch(zp_msm_workspace &ws, const BatchPlan &P, hipStream_t stream, uint64_t *h_out) {
  const size_t count = P.jobs.size();
  const HostJob *jobs = P.jobs.data();
  const bool sorted_path = P.sorted_path, msd_path = P.msd_path;
  const uint64_t K = P.K, points = P.points, counters = P.counters, n_max = P.n_max, GT = P.GT;
  const uint64_t max_units = P.max_units;
  const uint32_t max_nb = P.max_nb, max_lbits = P.max_lbits, max_groups = P.max_groups, max_chunks = P.max_chunks;
  const uint32_t Wt = P.Wt, log_u = P.log_u, log_g = P.log_g, gs = P.gs, S = P.S;
  const std::vector<SegDev> &segs = P.segs;
  const BatchDev &B = P.B;
  const FConsts &KC = consts();
  // ---- kernels: counting sort of (bucket, point) entries
  if (sorted_path) {
    if (points) {
      hipLaunchKernelGGL(k_digits_make, dim3(grid_for(points)), dim3(MSM_THREADS), 0, stream, B, ws.digits);
      ZP_HIP(hipGetLastError());
      hipLaunchKernelGGL(k_seg_count, dim3((unsigned)segs.size()), dim3(1024), max_nb * sizeof(uint32_t), stream, B,
                         ws.segs, ws.digits, ws.counts);
      ZP_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(MSM_THREADS), 0, stream, ws.counts + counters, (uint64_t)1, 0u);
    if (!points) {
      hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(counters + 1)), dim3(MSM_THREADS), 0, stream, ws.counts,
                         (uint64_t)counters + 1, 0u);
    }
    ZP_HIP(hipGetLastError());
    size_t tb0 = ws.cap_scan;
    ZP_HIP(rocprim::exclusive_scan(ws.scan_tmp, tb0, ws.counts, ws.scanned, 0u, (size_t)counters + 1,
                                   rocprim::plus<uint32_t>(), stream));
    hipLaunchKernelGGL(k_bucket_offsets, dim3(grid_for(K + 1)), dim3(MSM_THREADS), 0, stream, B, ws.scanned,
                       (uint64_t)counters, ws.off);
    hipLaunchKernelGGL(k_bucket_counts_from_off, dim3(grid_for(K)), dim3(MSM_THREADS), 0, stream, ws.off,
                       (uint32_t)K, ws.cnt);
    ZP_HIP(hipGetLastError());
    if (points) {
      hipLaunchKernelGGL(k_seg_scatter, dim3((unsigned)segs.size()), dim3(1024), max_nb * sizeof(uint32_t), stream,
                         B, ws.segs, ws.digits, ws.scanned, ws.vals);
      ZP_HIP(hipGetLastError());
    }
  } else if (msd_path) {
    const size_t lds_a = (size_t)max_groups * sizeof(uint32_t);
    const size_t lds_sc = msd_scatter_lds_words(max_groups) * sizeof(uint32_t);
    // pass B: cursors and tile ends per bucket, scan scratch, block tables,
    // the chunk offsets, then as large a tile as the LDS holds
    const size_t fixed_b = msd_group_fixed_words(max_lbits, max_chunks);
    const uint32_t tile_b = (uint32_t)std::min<size_t>(MSD_B_PER * 1024, ((MSD_LDS_WORDS - fixed_b) / 1024) * 1024);
    const size_t lds_b = (fixed_b + tile_b) * sizeof(uint32_t);
    if (points) {
      hipLaunchKernelGGL(k_digits_make, dim3(grid_for(points)), dim3(MSM_THREADS), 0, stream, B, ws.digits);
      ZP_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_msd_count, dim3((unsigned)segs.size()), dim3(1024), lds_a, stream, B, ws.segs, ws.digits,
                       ws.counts);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(MSM_THREADS), 0, stream, ws.counts + counters, (uint64_t)1, 0u);
    ZP_HIP(hipGetLastError());
    size_t tb0 = ws.cap_scan;
    ZP_HIP(rocprim::exclusive_scan(ws.scan_tmp, tb0, ws.counts, ws.scanned, 0u, (size_t)counters + 1,
                                   rocprim::plus<uint32_t>(), stream));
    if (points) {
      hipLaunchKernelGGL(k_msd_scatter, dim3((unsigned)segs.size()), dim3(1024), lds_sc, stream, B, ws.segs,
                         ws.digits, ws.scanned, ws.tmp, ws.vals);
      ZP_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_msd_group, dim3((unsigned)GT), dim3(1024), lds_b, stream, B, ws.win_g0, Wt, (uint32_t)GT,
                       max_lbits, max_chunks, tile_b, ws.scanned, ws.tmp, ws.vals, ws.off, ws.cnt);
    ZP_HIP(hipGetLastError());
  } else {
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(K + 1)), dim3(MSM_THREADS), 0, stream, ws.cnt, (uint64_t)K + 1, 0u);
    ZP_HIP(hipGetLastError());
    if (points) {
      hipLaunchKernelGGL(k_digits<0>, dim3(grid_for(points)), dim3(MSM_THREADS), 0, stream, B, ws.cnt, nullptr,
                         nullptr, nullptr);
      ZP_HIP(hipGetLastError());
    }
    size_t tb = ws.cap_scan;
    ZP_HIP(rocprim::exclusive_scan(ws.scan_tmp, tb, ws.cnt, ws.off, 0u, (size_t)K + 1, rocprim::plus<uint32_t>(),
                                   stream));
    ZP_HIP(hipMemcpyAsync(ws.cursor, ws.off, (K + 1) * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream));
    if (points) {
      hipLaunchKernelGGL(k_digits<1>, dim3(grid_for(points)), dim3(MSM_THREADS), 0, stream, B, nullptr, ws.cursor,
                         ws.vals, nullptr);
      ZP_HIP(hipGetLastError());
    }
  }
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(MSM_THREADS), 0, stream, ws.umax, (uint64_t)1, 0u);
  hipLaunchKernelGGL(k_units_per_bucket, dim3(grid_for(K + 1)), dim3(MSM_THREADS), 0, stream, ws.off, ws.cnt,
                     (uint32_t)K, log_u, ws.nunits, ws.umax);
  ZP_HIP(hipGetLastError());
  size_t tb = ws.cap_scan;
  ZP_HIP(rocprim::exclusive_scan(ws.scan_tmp, tb, ws.nunits, ws.ustart, 0u, (size_t)K + 1,
                                 rocprim::plus<uint32_t>(), stream));
  const uint32_t *total_units = ws.ustart + K;  // the scan's last entry
  hipLaunchKernelGGL(k_fill_units, dim3(grid_for(max_units)), dim3(MSM_THREADS), 0, stream, ws.ustart, (uint32_t)K,
                     total_units, ws.ubucket);
  ZP_HIP(hipGetLastError());
  const uint64_t slices = (std::max<uint64_t>(P.entries, 1) + (1ull << log_u) - 1) >> log_u;
  hipLaunchKernelGGL(k_accumulate, dim3(grid_for(slices, MSM_THREADS, 1u << 20)), dim3(MSM_THREADS), 0, stream, B,
                     ws.vals, ws.off, ws.ustart, (uint32_t)K, log_u, ws.partial, KC);
  ZP_HIP(hipGetLastError());
  // A bucket holds at most n_max entries (one digit per point and window),
  // so it touches at most ceil(n_max / U) + 1 slices: that many tree levels
  // suffice.
  const uint64_t units_per_bucket = ((n_max + (1ull << log_u) - 1) >> log_u) + 1;
  for (uint64_t stride = 1; stride < units_per_bucket; stride <<= 1) {
    // (a level with no work returns at once; 2048 blocks keep that cheap)
    hipLaunchKernelGGL(k_unit_tree, dim3(grid_for(max_units, MSM_THREADS, 2048)), dim3(MSM_THREADS), 0, stream,
                       ws.partial, ws.ubucket,
                       ws.ustart, ws.nunits, total_units, ws.umax, (uint32_t)stride, KC);
    ZP_HIP(hipGetLastError());
  }
  ZP_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_window_groups, dim3(grid_for(gs)), dim3(MSM_THREADS), 0, stream, ws.partial, ws.ustart, ws.cnt,
                     ws.wins,
                     ws.gstart, Wt, log_g, ws.groups, KC);
  ZP_HIP(hipGetLastError());
  // Windows of many groups (a shifted-table job's one window) are summed in
  // slices, then the slices per window.
  if (S > 1) {
    hipLaunchKernelGGL(k_window_sum, dim3(Wt, S), dim3(MSM_THREADS), MSM_THREADS * sizeof(Xyzt), stream, ws.groups,
                       ws.gstart, ws.winpart, KC);
    ZP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_window_slices, dim3(grid_for(Wt, 64)), dim3(64), 0, stream, ws.winpart, Wt, S, ws.winsum, KC);
  } else {
    hipLaunchKernelGGL(k_window_sum, dim3(Wt), dim3(MSM_THREADS), MSM_THREADS * sizeof(Xyzt), stream, ws.groups,
                       ws.gstart, ws.winsum, KC);
  }
  ZP_HIP(hipGetLastError());
  // exceptional-point scalar sums, per job and group
  for (size_t j = 0; j < count; j++) {
    const zp_msm_context *c = jobs[j].ctx;
    if (c->nc_groups.empty()) continue;
    hipLaunchKernelGGL(k_nc_sums, dim3((unsigned)c->nc_groups.size()), dim3(256), 0, stream, jobs[j].d_scalars,
                       c->d_nc_idx, c->d_nc_off, (uint64_t)jobs[j].n, ws.nc_sums + 5 * P.nc0[j]);
    ZP_HIP(hipGetLastError());
  }
  if (P.host_fin) return ZP_OK;
  // device finalisation: window combine, exceptional points, wire form
  hipLaunchKernelGGL(k_msm_finalize, dim3((unsigned)count), dim3(4), 0, stream, ws.fin, ws.wins, ws.winsum,
                     ws.nc_sums, ws.d_out, ws.d_consts);
  ZP_HIP(hipGetLastError());
  if (h_out)
    ZP_HIP(hipMemcpyAsync(h_out, ws.d_out, count * 18 * sizeof(uint64_t), hipMemcpyDeviceToHost, stream));
  return ZP_OK;
}

// Plans, launches and waits for one batch; writes every job's out.
static int run_batch(zp_msm_workspace &ws, const HostJob *jobs, size_t count, const MsmOptions &bopts,
                     hipStream_t stream) {
  if (count == 0) return ZP_OK;
  BatchPlan P;
  const int rc = plan_batch(ws, jobs, count, bopts, stream, P);
  if (rc != ZP_OK) {
    (void)hipStreamSynchronize(stream);  // table uploads already issued read P's vectors
    return rc;
  }
  if (!P.host_fin) {
    ws.h_out.resize(18 * count);
    ZP_TRY(launch_batch(ws, P, stream, ws.h_out.data()));
    ZP_HIP(hipStreamSynchronize(stream));
    for (size_t j = 0; j < count; j++) memcpy(jobs[j].out, ws.h_out.data() + 18 * j, sizeof(g1_projective_t));
    return ZP_OK;
  }
  ZP_TRY(launch_batch(ws, P, stream, nullptr));
  const std::vector<size_t> &nc0 = P.nc0;
  ws.h_winsum.resize(P.Wt);
  ZP_HIP(hipMemcpyAsync(ws.h_winsum.data(), ws.winsum, P.Wt * sizeof(Xyzt), hipMemcpyDeviceToHost, stream));
  if (nc0[count]) {
    ws.h_nc_sums.resize(5 * nc0[count]);
    ZP_HIP(hipMemcpyAsync(ws.h_nc_sums.data(), ws.nc_sums, 5 * nc0[count] * sizeof(uint64_t), hipMemcpyDeviceToHost,
                          stream));
  }
  ZP_HIP(hipStreamSynchronize(stream));
  // ---- host finalisation (window combine, conversion, exceptional points)
  auto fin = [&](size_t j) {
    finalize(P.layouts[j], ws.h_winsum.data() + P.jd[j].win0, jobs[j].ctx->nc_groups,
             ws.h_nc_sums.data() + 5 * nc0[j], jobs[j].out);
  };
  const size_t nthreads = std::min<size_t>(count, 16);
  if (nthreads <= 1) {
    for (size_t j = 0; j < count; j++) fin(j);
  } else {
    std::vector<std::thread> th;
    std::atomic<size_t> next{0};
    for (size_t t = 0; t < nthreads; t++)
      th.emplace_back([&] {
        for (size_t j = next++; j < count; j = next++) fin(j);
      });
    for (auto &t : th) t.join();
  }
  return ZP_OK;
}

// Digit dump for one job (diagnostics).
static int dump_digits(const MsmOptions &o, const uint64_t *d_scalars, size_t n, int32_t *d_out, hipStream_t s,
                       int *W_out) {
  HostLayout L = choose_layout(o, n);
  *W_out = L.W;
  if (!d_out || !n) return ZP_OK;
  MsmJobDev jd{};
  jd.scalars = d_scalars;
  jd.n = (uint32_t)n;
  jd.W = (uint32_t)L.W;
  std::vector<WinDev> wins(L.W);
  uint32_t base = 0;
  for (int w = 0; w < L.W; w++) {
    wins[w] = WinDev{0, (uint32_t)L.off[w], (uint32_t)L.width[w], L.nb[w], base, (uint32_t)(w == L.W - 1)};
    base += L.nb[w];
  }
  MsmJobDev *d_jobs = nullptr;
  WinDev *d_wins = nullptr;
  uint32_t *d_zero = nullptr;
  ZP_HIP(hipMalloc(&d_jobs, sizeof(jd)));
  ZP_HIP(hipMalloc(&d_wins, wins.size() * sizeof(WinDev)));
  ZP_HIP(hipMalloc(&d_zero, 4));
  ZP_HIP(hipMemcpy(d_jobs, &jd, sizeof(jd), hipMemcpyHostToDevice));
  ZP_HIP(hipMemcpy(d_wins, wins.data(), wins.size() * sizeof(WinDev), hipMemcpyHostToDevice));
  ZP_HIP(hipMemset(d_zero, 0, 4));
  BatchDev B{};
  B.jobs = d_jobs;
  B.job_pt0 = d_zero;
  B.job_b0 = d_zero;
  B.wins = d_wins;
  B.J = 1;
  B.total_points = (uint32_t)n;
  B.K = base;
  hipLaunchKernelGGL(k_digits<2>, dim3(grid_for(n)), dim3(MSM_THREADS), 0, s, B, nullptr, nullptr, nullptr, d_out);
  ZP_HIP(hipGetLastError());
  ZP_HIP(hipStreamSynchronize(s));
  (void)hipFree(d_jobs);
  (void)hipFree(d_wins);
  (void)hipFree(d_zero);
  return ZP_OK;
}

}  // namespace zp

using namespace zp;

// -------------------------------------------- harness data-file scalars --
__global__ void k_fr_to_montgomery(biginteger256_t *s, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    Fr8 v;
    const uint32_t *w = reinterpret_cast<const uint32_t *>(s[i].data);
#pragma unroll
    for (int j = 0; j < 8; j++) v.w[j] = w[j];
    const Fr8 m = fr_to_montgomery(v);
    uint32_t *o = reinterpret_cast<uint32_t *>(s[i].data);
#pragma unroll
    for (int j = 0; j < 8; j++) o[j] = m.w[j];
  }
}

// ---------------------------------------------------------------- init --
static int init_points(zp_msm_context *ctx, const g1_affine_t *points, bool on_device) {
  const FConsts &K = consts();
  const F to_r = ctx->opts.canonical_coords ? K.r2 : K.to_int;
  const size_t n = ctx->n;
  ZP_HIP(hipMalloc(&ctx->d_pts, std::max<size_t>(n, 1) * sizeof(PackedPre)));
  ZP_HIP(hipMalloc(&ctx->d_conv, std::max<size_t>(n, 1)));
  const size_t SLICE = std::min<size_t>(n ? n : 1, (size_t)1 << 21);
  g1_affine_t *d_stage = nullptr;
  F *d_u = nullptr, *d_v = nullptr, *d_den = nullptr, *d_tmp = nullptr;
  uint8_t *d_flag = nullptr;
  int st = ZP_OK;
  auto fail = [&](const char *what) {
    set_last_error("msm_init: %s", what);
    st = ZP_ERR_OUT_OF_MEMORY;
  };
  if (!on_device && hipMalloc(&d_stage, SLICE * sizeof(g1_affine_t)) != hipSuccess) fail("staging allocation");
  if (st == ZP_OK && (hipMalloc(&d_u, SLICE * sizeof(F)) != hipSuccess || hipMalloc(&d_v, SLICE * sizeof(F)) != hipSuccess ||
                      hipMalloc(&d_den, SLICE * sizeof(F)) != hipSuccess ||
                      hipMalloc(&d_tmp, SLICE * sizeof(F)) != hipSuccess || hipMalloc(&d_flag, SLICE) != hipSuccess))
    fail("temporary allocation");
  hipStream_t s = ctx->stream;
  for (size_t b = 0; st == ZP_OK && b < n; b += SLICE) {
    const size_t m = std::min(SLICE, n - b);
    const g1_affine_t *src = points + b;
    if (!on_device) {
      if (hipMemcpyAsync(d_stage, points + b, m * sizeof(g1_affine_t), hipMemcpyHostToDevice, s) != hipSuccess) {
        set_last_error("msm_init: point upload failed");
        st = ZP_ERR_HIP;
        break;
      }
      src = d_stage;
    }
    hipLaunchKernelGGL(k_pre_a, dim3(grid_for(m)), dim3(MSM_THREADS), 0, s, src, (uint64_t)m, K, to_r, d_u, d_v,
                       d_den, d_flag);
    const uint32_t chunk = 32;
    hipLaunchKernelGGL(k_batch_inv, dim3(grid_for((m + chunk - 1) / chunk, 64)), dim3(64), 0, s, d_den, (uint64_t)m,
                       chunk, d_tmp, K);
    hipLaunchKernelGGL(k_pre_c, dim3(grid_for(m)), dim3(MSM_THREADS), 0, s, d_u, d_v, d_den, d_flag, (uint64_t)m, K,
                       ctx->d_pts + b, ctx->d_conv + b);
    if (hipGetLastError() != hipSuccess) {
      set_last_error("msm_init: precompute launch failed");
      st = ZP_ERR_HIP;
    }
  }
  if (st == ZP_OK && hipStreamSynchronize(s) != hipSuccess) {
    set_last_error("msm_init: precompute failed");
    st = ZP_ERR_HIP;
  }
  if (d_stage) (void)hipFree(d_stage);
  if (d_u) (void)hipFree(d_u);
  if (d_v) (void)hipFree(d_v);
  if (d_den) (void)hipFree(d_den);
  if (d_tmp) (void)hipFree(d_tmp);
  if (d_flag) (void)hipFree(d_flag);
  if (st != ZP_OK) return st;
  // Non-convertible points (driver.cpp:160-172), grouped by value.
  std::vector<uint8_t> conv(n);
  if (n) ZP_HIP(hipMemcpy(conv.data(), ctx->d_conv, n, hipMemcpyDeviceToHost));
  for (size_t i = 0; i < n; i++) {
    if (conv[i]) continue;
    g1_affine_t p;
    if (on_device)
      ZP_HIP(hipMemcpy(&p, points + i, sizeof(p), hipMemcpyDeviceToHost));
    else
      p = points[i];
    U384 x, y;
    memcpy(x.w, p.x.data, 48);
    memcpy(y.w, p.y.data, 48);
    const F fx = f_canonical(f_mul(f_from_u384(x), to_r));
    const F fy = f_canonical(f_mul(f_from_u384(y), to_r));
    NcGroup *grp = nullptr;
    for (NcGroup &g : ctx->nc_groups)
      if (f_eq(g.x, fx) && f_eq(g.y, fy)) grp = &g;
    if (!grp) {
      ctx->nc_groups.push_back(NcGroup{fx, fy, {}});
      grp = &ctx->nc_groups.back();
    }
    grp->idx.push_back((uint32_t)i);
    ctx->num_nonconv++;
  }
  if (!ctx->nc_groups.empty()) {
    std::vector<uint32_t> all, off{0};
    for (const NcGroup &g : ctx->nc_groups) {
      all.insert(all.end(), g.idx.begin(), g.idx.end());
      off.push_back((uint32_t)all.size());
    }
    ZP_HIP(hipMalloc(&ctx->d_nc_idx, all.size() * sizeof(uint32_t)));
    ZP_HIP(hipMalloc(&ctx->d_nc_off, off.size() * sizeof(uint32_t)));
    ZP_HIP(hipMemcpy(ctx->d_nc_idx, all.data(), all.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    ZP_HIP(hipMemcpy(ctx->d_nc_off, off.data(), off.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    // P, 2P, 3P of each group for k_msm_finalize.  A curve point without a
    // twisted-Edwards image has order 2 or 4, so S P = (S mod 4) P; anything
    // else is not a curve point and is rejected.
    std::vector<NcMultDev> mult(ctx->nc_groups.size());
    for (size_t g = 0; g < ctx->nc_groups.size(); g++) {
      const NcGroup &grp = ctx->nc_groups[g];
      Jac m;
      m.X = grp.x;
      m.Y = grp.y;
      m.Z = K.one;
      for (int k = 0; k < 4; k++) {
        if (k > 0) jac_add_aff(m, grp.x, grp.y, K);  // m = (k + 1) P
        if (k == 3) break;
        NcMultDev &d = mult[g];
        d.inf[k] = f_is_zero_canon(m.Z);
        d.x[k] = f_zero();
        d.y[k] = f_zero();
        if (!d.inf[k]) {
          const F zi = f_inv(m.Z, K.one), zi2 = f_sqr(zi);
          d.x[k] = f_canonical(f_mul(m.X, zi2));
          d.y[k] = f_canonical(f_mul(m.Y, f_mul(zi2, zi)));
        }
      }
      ZP_CHECK(f_is_zero_canon(m.Z), ZP_ERR_INVALID_ARG,
               "msm_init: point %u has no twisted-Edwards image and is not a curve point (4 P != 0)",
               (unsigned)grp.idx[0]);
    }
    ZP_HIP(hipMalloc(&ctx->d_nc_mult, mult.size() * sizeof(NcMultDev)));
    ZP_HIP(hipMemcpy(ctx->d_nc_mult, mult.data(), mult.size() * sizeof(NcMultDev), hipMemcpyHostToDevice));
  }
  return ZP_OK;
}

// Width of the window-shifted table for n points ("precompute=auto"): the
// cheapest (msm_cost, shared buckets) among the widths whose top window is
// nearly full -- 253 = 10 x 23 + 23 = 14 x 17 + 15 = 15 x 16 + 13 =
// 16 x 15 + 13 = 22 x 11 + 11.  A short top window (22 bits: 11 x 22 + 11)
// puts every point's top digit into the few lowest buckets, i.e. into one or
// two groups of the bucket sort, and the block sorting them serializes
// (measured at 2^24: 27 ms against 1.5 ms).  2^26 points: 23 bits (11 tables).
static int auto_shift_width(size_t n) {
  static const int widths[] = {23, 17, 16, 15, 11};
  int best = 11;
  double best_cost = 0;
  for (int c : widths) {
    HostLayout L = layout_uniform(c);
    layout_finish(L);
    const double cost = msm_cost(n, L, true);
    if (best_cost == 0 || cost < best_cost) {
      best = c;
      best_cost = cost;
    }
  }
  return best;
}

// The window-shifted table for the uniform layout of c bits (see k_shift_a).
// Slices of points bound the temporaries; entry w of a slice is made from
// its entry w - 1, window after window.
static int init_shift_table(zp_msm_context *ctx, int c) {
  const FConsts &K = consts();
  const size_t n = ctx->n;
  HostLayout L = layout_uniform(c);
  const uint32_t W = (uint32_t)L.W;
  ZP_CHECK((uint64_t)n * W < (1ull << 31), ZP_ERR_TOO_LARGE, "msm_init: precompute: %zu points x %u windows >= 2^31",
           n, W);
  if (n == 0) return ZP_OK;
  size_t free_b = 0, total_b = 0;
  ZP_HIP(hipMemGetInfo(&free_b, &total_b));
  const size_t bytes = n * W * sizeof(PackedPre);
  // and room for one full-table MSM's workspace: per (point, window) a digit,
  // a pass-A entry and a sorted entry, the unit partials and the shared
  // buckets, plus slack
  const size_t ws_bytes = n * W * 16 + (size_t(1) << (c - 1)) * 3 * sizeof(Xyzt) + (size_t(2) << 30);
  ZP_CHECK(bytes + ws_bytes < free_b, ZP_ERR_OUT_OF_MEMORY,
           "msm_init: precompute=%d needs %.1f GB for %u shifted tables (+%.1f GB of MSM workspace), %.1f GB free",
           c, bytes / 1e9, W, ws_bytes / 1e9, free_b / 1e9);
  ZP_HIP(hipMalloc(&ctx->d_shift, bytes));
  const size_t SLICE = std::min<size_t>(n, (size_t)1 << 22);
  F *X = nullptr, *Y = nullptr, *Z = nullptr, *T = nullptr;
  int *d_bad = nullptr;
  int st = ZP_OK;
  if (hipMalloc(&X, SLICE * sizeof(F)) != hipSuccess || hipMalloc(&Y, SLICE * sizeof(F)) != hipSuccess ||
      hipMalloc(&Z, SLICE * sizeof(F)) != hipSuccess || hipMalloc(&T, SLICE * sizeof(F)) != hipSuccess ||
      hipMalloc(&d_bad, sizeof(int)) != hipSuccess) {
    set_last_error("msm_init: precompute: temporary allocation failed");
    st = ZP_ERR_OUT_OF_MEMORY;
  }
  hipStream_t s = ctx->stream;
  if (st == ZP_OK) {
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(MSM_THREADS), 0, s, (uint32_t *)d_bad, (uint64_t)1, 0u);
    hipLaunchKernelGGL(k_shift_copy0, dim3(grid_for(n)), dim3(MSM_THREADS), 0, s, ctx->d_pts, (uint64_t)n, W,
                       ctx->d_shift);
    for (uint32_t w = 1; w < W; w++) {
      for (size_t b = 0; b < n; b += SLICE) {
        const size_t m = std::min(SLICE, n - b);
        hipLaunchKernelGGL(k_shift_a, dim3(grid_for(m)), dim3(MSM_THREADS), 0, s, ctx->d_shift, (uint64_t)b,
                           (uint64_t)m, W, w, (uint32_t)c, ctx->d_conv, K, X, Y, Z, d_bad);
        const uint32_t chunk = 32;
        hipLaunchKernelGGL(k_batch_inv, dim3(grid_for((m + chunk - 1) / chunk, 64)), dim3(64), 0, s, Z, (uint64_t)m,
                           chunk, T, K);
        hipLaunchKernelGGL(k_shift_c, dim3(grid_for(m)), dim3(MSM_THREADS), 0, s, X, Y, Z, (uint64_t)b, (uint64_t)m,
                           W, w, K, ctx->d_shift);
      }
    }
    if (hipGetLastError() != hipSuccess) {
      set_last_error("msm_init: precompute launch failed");
      st = ZP_ERR_HIP;
    }
  }
  int bad = 0;
  if (st == ZP_OK && (hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
                      hipStreamSynchronize(s) != hipSuccess)) {
    set_last_error("msm_init: precompute failed");
    st = ZP_ERR_HIP;
  }
  if (X) (void)hipFree(X);
  if (Y) (void)hipFree(Y);
  if (Z) (void)hipFree(Z);
  if (T) (void)hipFree(T);
  if (d_bad) (void)hipFree(d_bad);
  if (st != ZP_OK || bad) {
    // A table point of small order has a multiple without a twisted-Edwards
    // image: jobs on this table keep the per-window buckets.
    (void)hipFree(ctx->d_shift);
    ctx->d_shift = nullptr;
    return st;
  }
  ctx->shift_c = c;
  ctx->shift_W = (int)W;
  return ZP_OK;
}

static int msm_init_common(const char *options, size_t options_len, const g1_affine_t *points, size_t npoints,
                           bool on_device, zp_msm_context **out) {
  ZP_CHECK(out != nullptr, ZP_ERR_INVALID_ARG, "msm_init: out is NULL");
  *out = nullptr;
  ZP_CHECK(points != nullptr || npoints == 0, ZP_ERR_INVALID_ARG, "msm_init: points is NULL");
  ZP_CHECK(npoints < (1ull << 31), ZP_ERR_TOO_LARGE, "msm_init: %zu points > 2^31 - 1", npoints);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    set_last_error("msm_init: no HIP device");
    return ZP_ERR_NO_DEVICE;
  }
  const MsmOptions opts = parse_options(options, options_len);
  ZP_TRY(validate_options(opts, "msm_init"));
  zp_msm_context *ctx = new zp_msm_context();
  ctx->n = npoints;
  ctx->opts = opts;
  int st = ZP_OK;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    set_last_error("msm_init: stream creation failed");
    st = ZP_ERR_HIP;
  }
  if (st == ZP_OK) st = init_points(ctx, points, on_device);
  if (st == ZP_OK && opts.precompute)
    st = init_shift_table(ctx, opts.precompute > 0 ? opts.precompute : auto_shift_width(npoints));
  if (st != ZP_OK) {
    msm_free(ctx);
    return st;
  }
  *out = ctx;
  return ZP_OK;
}

extern "C" int zp_msm_init(const char *options, size_t options_len, const g1_affine_t *points, size_t npoints,
                           zp_msm_context **out) {
  return msm_init_common(options, options_len, points, npoints, false, out);
}

extern "C" int zp_msm_init_device(const char *options, size_t options_len, const g1_affine_t *d_points,
                                  size_t npoints, zp_msm_context **out) {
  return msm_init_common(options, options_len, d_points, npoints, true, out);
}

extern "C" zp_msm_context *msm_init(const char *xclbin, ssize_t xclbin_len, g1_affine_t *rust_points,
                                    ssize_t npoints) {
  zp_msm_context *ctx = nullptr;
  if (npoints < 0 || xclbin_len < 0) {
    set_last_error("msm_init: negative length");
    return nullptr;
  }
  if (zp_msm_init(xclbin, (size_t)xclbin_len, rust_points, (size_t)npoints, &ctx) != ZP_OK) {
    fprintf(stderr, "msm_init failed: %s\n", zp_last_error());
    return nullptr;
  }
  return ctx;
}

extern "C" void msm_free(zp_msm_context *ctx) {
  if (!ctx) return;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  ctx->ws.release();
  if (ctx->d_scalars) (void)hipFree(ctx->d_scalars);
  if (ctx->d_scalars2) (void)hipFree(ctx->d_scalars2);
  ctx->uploader.release();
  if (ctx->d_pts) (void)hipFree(ctx->d_pts);
  if (ctx->d_shift) (void)hipFree(ctx->d_shift);
  if (ctx->d_conv) (void)hipFree(ctx->d_conv);
  if (ctx->d_nc_idx) (void)hipFree(ctx->d_nc_idx);
  if (ctx->d_nc_off) (void)hipFree(ctx->d_nc_off);
  if (ctx->d_nc_mult) (void)hipFree(ctx->d_nc_mult);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

extern "C" int zp_msm_mult(zp_msm_context *ctx, g1_projective_t *out, uint64_t num_batches,
                           const biginteger256_t *scalars) {
  ZP_CHECK(ctx != nullptr && out != nullptr && (scalars != nullptr || num_batches == 0 || ctx->n == 0),
           ZP_ERR_INVALID_ARG, "zp_msm_mult: NULL argument");
  const size_t n = ctx->n;
  ZP_TRY(grow(ctx->d_scalars, ctx->cap_scalars, std::max<size_t>(4 * n, 4)));
  if (num_batches > 1 && n) ZP_TRY(grow(ctx->d_scalars2, ctx->cap_scalars2, std::max<size_t>(4 * n, 4)));
  // The previous call's work on ctx->stream may still read d_scalars.
  if (n && num_batches) {
    ZP_HIP(hipStreamSynchronize(ctx->stream));
    ZP_TRY(ctx->uploader.upload(ctx->d_scalars, scalars, n * 32));
  }
  // Batches alternate between two device buffers: while batch b computes on
  // ctx->stream, a helper thread uploads batch b + 1 through the pinned
  // staging (driver.cpp:495-600 overlaps transfers with work across batches
  // the same way).  run_batch returns only when batch b is finished, so the
  // buffer the next upload overwrites is free.
  for (uint64_t b = 0; b < num_batches; b++) {
    uint64_t *cur = (b & 1) ? ctx->d_scalars2 : ctx->d_scalars;
    uint64_t *nxt = (b & 1) ? ctx->d_scalars : ctx->d_scalars2;
    int copy_rc = ZP_OK;
    std::thread copier;
    if (b + 1 < num_batches && n) {
      const biginteger256_t *hs = scalars + (b + 1) * n;
      int dev = 0;
      ZP_HIP(hipGetDevice(&dev));
      copier = std::thread([&, hs, nxt, dev] {
        if (hipSetDevice(dev) != hipSuccess || ctx->uploader.upload(nxt, hs, n * 32) != ZP_OK) copy_rc = ZP_ERR_HIP;
      });
    }
    HostJob job{ctx, n, cur, out + b};
    const int rc = run_batch(ctx->ws, &job, 1, MsmOptions(), ctx->stream);
    if (copier.joinable()) copier.join();
    ZP_TRY(rc);
    ZP_CHECK(copy_rc == ZP_OK, copy_rc, "zp_msm_mult: scalar upload of batch %llu failed",
             (unsigned long long)(b + 1));
  }
  return ZP_OK;
}

extern "C" void msm_mult(zp_msm_context *ctx, g1_projective_t *out, uint64_t num_batches,
                         biginteger256_t *ptr_scalars) {
  if (zp_msm_mult(ctx, out, num_batches, ptr_scalars) != ZP_OK)
    fprintf(stderr, "msm_mult failed: %s\n", zp_last_error());
}

extern "C" int zp_msm_mult_device(zp_msm_context *ctx, g1_projective_t *out, uint64_t num_batches,
                                  const biginteger256_t *d_scalars, void *stream) {
  ZP_CHECK(ctx != nullptr && out != nullptr && (d_scalars != nullptr || num_batches == 0 || ctx->n == 0),
           ZP_ERR_INVALID_ARG, "zp_msm_mult_device: NULL argument");
  hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
  for (uint64_t b = 0; b < num_batches; b++) {
    HostJob job{ctx, ctx->n, reinterpret_cast<const uint64_t *>(d_scalars + b * ctx->n), out + b};
    ZP_TRY(run_batch(ctx->ws, &job, 1, MsmOptions(), s));
  }
  return ZP_OK;
}

extern "C" int zp_msm_mult_prefix_device(zp_msm_context *ctx, size_t n, g1_projective_t *out,
                                         const biginteger256_t *d_scalars, void *stream) {
  ZP_CHECK(ctx != nullptr && out != nullptr, ZP_ERR_INVALID_ARG, "zp_msm_mult_prefix_device: NULL argument");
  ZP_CHECK(n <= ctx->n, ZP_ERR_TOO_LARGE, "zp_msm_mult_prefix_device: n %zu > %zu points", n, ctx->n);
  ZP_CHECK(d_scalars != nullptr || n == 0, ZP_ERR_INVALID_ARG, "zp_msm_mult_prefix_device: NULL scalars");
  hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
  HostJob job{ctx, n, reinterpret_cast<const uint64_t *>(d_scalars), out};
  return run_batch(ctx->ws, &job, 1, MsmOptions(), s);
}

extern "C" int zp_fr_to_montgomery_device(biginteger256_t *d_scalars, uint64_t n, void *stream) {
  ZP_CHECK(d_scalars != nullptr || n == 0, ZP_ERR_INVALID_ARG, "zp_fr_to_montgomery_device: NULL scalars");
  if (n == 0) return ZP_OK;
  hipLaunchKernelGGL(k_fr_to_montgomery, dim3(grid_for(n)), dim3(MSM_THREADS), 0, (hipStream_t)stream, d_scalars, n);
  ZP_HIP(hipGetLastError());
  return ZP_OK;
}

// ------------------------------------------------------------- batched --
extern "C" int zp_msm_workspace_create(zp_msm_workspace **out) {
  ZP_CHECK(out != nullptr, ZP_ERR_INVALID_ARG, "zp_msm_workspace_create: NULL out");
  *out = new zp_msm_workspace();
  return ZP_OK;
}

extern "C" int zp_msm_workspace_destroy(zp_msm_workspace *ws) {
  if (ws) {
    ws->release();
    delete ws;
  }
  return ZP_OK;
}

extern "C" int zp_msm_batch_device(zp_msm_workspace *ws, const zp_msm_job *jobs, size_t count, const char *options,
                                   size_t options_len, void *stream) {
  ZP_CHECK(ws != nullptr && (jobs != nullptr || count == 0), ZP_ERR_INVALID_ARG, "zp_msm_batch_device: NULL argument");
  const MsmOptions o = parse_options(options, options_len);
  ZP_TRY(validate_options(o, "zp_msm_batch_device"));
  std::vector<HostJob> hj(count);
  for (size_t j = 0; j < count; j++) {
    const zp_msm_job &J = jobs[j];
    ZP_CHECK(J.ctx != nullptr && J.out != nullptr, ZP_ERR_INVALID_ARG, "job %zu: NULL context or output", j);
    ZP_CHECK(J.n <= J.ctx->n, ZP_ERR_TOO_LARGE, "job %zu: n %zu > %zu points", j, J.n, J.ctx->n);
    ZP_CHECK(J.d_scalars != nullptr || J.n == 0, ZP_ERR_INVALID_ARG, "job %zu: NULL scalars", j);
    ZP_CHECK(J.window == 0 || (J.window >= 8 && J.window <= 24), ZP_ERR_INVALID_ARG, "job %zu: window %d", j,
             J.window);
    hj[j] = HostJob{J.ctx, J.n, reinterpret_cast<const uint64_t *>(J.d_scalars), J.out, J.window};
  }
  return run_batch(*ws, hj.data(), count, o, (hipStream_t)stream);
}

// ---------------------------------------------- allocation-free entry --
// Device working memory of zp_msm_mult_prefix_noalloc: one array per GPU,
// held until the process ends.  Host code never uses the array's name as a
// device address; it takes the address once per device from
// hipGetSymbolAddress and passes that pointer on.
static constexpr size_t ZP_MSM_SCRATCH_BYTES = size_t(64) << 20;
static __device__ __align__(256) char scratch[ZP_MSM_SCRATCH_BYTES];

static constexpr int ZP_MAX_DEVICES = 64;
static std::mutex g_noalloc_mu[ZP_MAX_DEVICES];
static char *g_scratch_ptr[ZP_MAX_DEVICES];            // per device, from hipGetSymbolAddress
static zp_msm_workspace g_noalloc_ws[ZP_MAX_DEVICES];  // arena-mode workspace over that array

extern "C" int zp_msm_mult_prefix_noalloc(zp_msm_context *ctx, size_t n, g1_projective_t *out,
                                          const biginteger256_t *d_scalars, void *stream) {
  ZP_CHECK(ctx != nullptr && out != nullptr, ZP_ERR_INVALID_ARG, "zp_msm_mult_prefix_noalloc: NULL argument");
  ZP_CHECK(n <= ctx->n, ZP_ERR_TOO_LARGE, "zp_msm_mult_prefix_noalloc: n %zu > %zu points", n, ctx->n);
  ZP_CHECK(d_scalars != nullptr || n == 0, ZP_ERR_INVALID_ARG, "zp_msm_mult_prefix_noalloc: NULL scalars");
  int dev = 0;
  ZP_HIP(hipGetDevice(&dev));
  ZP_CHECK(dev >= 0 && dev < ZP_MAX_DEVICES, ZP_ERR_INVALID_ARG, "zp_msm_mult_prefix_noalloc: device %d", dev);
  // One call at a time per GPU: every call on a device shares its scratch array.
  std::lock_guard<std::mutex> lock(g_noalloc_mu[dev]);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    ZP_CHECK(e == hipSuccess && ptr != nullptr, ZP_ERR_HIP,
             "zp_msm_mult_prefix_noalloc: scratch symbol address on device %d: %s", dev, hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<char *>(ptr);
  }
  zp_msm_workspace &ws = g_noalloc_ws[dev];
  ws.arena = g_scratch_ptr[dev];
  ws.arena_cap = ZP_MSM_SCRATCH_BYTES;
  ws.release();  // arena mode: forget the last call's buffers, carve again from the start
  hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
  HostJob job{ctx, n, reinterpret_cast<const uint64_t *>(d_scalars), out};
  BatchPlan P;
  const int rc = plan_batch(ws, &job, 1, MsmOptions(), s, P);
  if (rc != ZP_OK) {
    (void)hipStreamSynchronize(s);  // table uploads already issued read P's vectors
    return rc;
  }
  P.host_fin = false;  // results come from k_msm_finalize straight into out
  ZP_TRY(launch_batch(ws, P, s, reinterpret_cast<uint64_t *>(out)));
  ZP_HIP(hipStreamSynchronize(s));
  return ZP_OK;
}

// ---------------------------------------------------------- transfers --
extern "C" int zp_transfer_create(zp_transfer **out) {
  ZP_CHECK(out != nullptr, ZP_ERR_INVALID_ARG, "zp_transfer_create: NULL out");
  *out = new zp_transfer();
  return ZP_OK;
}

extern "C" int zp_transfer_destroy(zp_transfer *t) {
  if (t) {
    t->u.release();
    delete t;
  }
  return ZP_OK;
}

extern "C" int zp_transfer_h2d(zp_transfer *t, void *d_dst, const void *h_src, size_t bytes) {
  ZP_CHECK(t != nullptr && (bytes == 0 || (d_dst && h_src)), ZP_ERR_INVALID_ARG, "zp_transfer_h2d: NULL argument");
  return t->u.upload(d_dst, h_src, bytes);
}

extern "C" int zp_transfer_d2h(zp_transfer *t, void *h_dst, const void *d_src, size_t bytes, void *after_stream) {
  ZP_CHECK(t != nullptr && (bytes == 0 || (h_dst && d_src)), ZP_ERR_INVALID_ARG, "zp_transfer_d2h: NULL argument");
  return t->u.download(h_dst, d_src, bytes, (hipStream_t)after_stream);
}

// ------------------------------------------------------- mixed batch --
// NTT groups and one MSM batch captured into ONE HIP graph: the NTT launches
// on the capture stream and the MSM launch sequence on a forked stream (two
// parallel branches of the graph), the MSM results copied into pinned host
// memory by the graph's last node.  Planning, allocation and table uploads
// happen once at creation; each launch replays every kernel from the
// current contents of the NTT vectors and the scalar buffers.
struct zp_mixed_batch {
  zp_ntt_batch *ntt = nullptr;
  zp_msm_workspace ws;
  BatchPlan plan;
  uint64_t *h_out = nullptr;  // pinned, 18 words per MSM job
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

extern "C" int zp_mixed_batch_destroy(zp_mixed_batch *b) {
  if (!b) return ZP_OK;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->ntt) (void)zp_ntt_batch_destroy(b->ntt);
  if (b->h_out) (void)hipHostFree(b->h_out);
  b->ws.release();
  delete b;
  return ZP_OK;
}

static int mixed_capture(zp_mixed_batch *b) {
  hipStream_t cap = nullptr, s2 = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  int st = ZP_OK;
  auto hip_ok = [&](hipError_t e, const char *what) {
    if (e != hipSuccess && st == ZP_OK) {
      set_last_error("zp_mixed_batch_create: %s: %s", what, hipGetErrorString(e));
      st = ZP_ERR_HIP;
    }
  };
  hip_ok(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking), "stream");
  hip_ok(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking), "stream");
  hip_ok(hipEventCreateWithFlags(&fork, hipEventDisableTiming), "event");
  hip_ok(hipEventCreateWithFlags(&join, hipEventDisableTiming), "event");
  if (st == ZP_OK) {
    hip_ok(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal), "begin capture");
    if (st == ZP_OK) {
      hip_ok(hipEventRecord(fork, cap), "fork");
      hip_ok(hipStreamWaitEvent(s2, fork, 0), "fork wait");
      if (st == ZP_OK && b->ntt) st = zp_ntt_batch_launch(b->ntt, 0, cap);
      if (st == ZP_OK && !b->plan.jobs.empty()) st = launch_batch(b->ws, b->plan, s2, b->h_out);
      hip_ok(hipEventRecord(join, s2), "join");
      hip_ok(hipStreamWaitEvent(cap, join, 0), "join wait");
      hipGraph_t g = nullptr;
      hip_ok(hipStreamEndCapture(cap, &g), "end capture");
      b->graph = g;
    }
  }
  if (st == ZP_OK) hip_ok(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "instantiate");
  if (fork) (void)hipEventDestroy(fork);
  if (join) (void)hipEventDestroy(join);
  if (s2) (void)hipStreamDestroy(s2);
  if (cap) (void)hipStreamDestroy(cap);
  return st;
}

extern "C" int zp_mixed_batch_create(uint64_t *const *ntt_ptrs, const unsigned *ntt_log_ns, const int *ntt_dirs,
                                     size_t ntt_count, const zp_msm_job *msm_jobs, size_t msm_count,
                                     const char *options, size_t options_len, zp_mixed_batch **out) {
  ZP_CHECK(out != nullptr, ZP_ERR_INVALID_ARG, "zp_mixed_batch_create: out is NULL");
  *out = nullptr;
  ZP_CHECK(msm_jobs != nullptr || msm_count == 0, ZP_ERR_INVALID_ARG, "zp_mixed_batch_create: NULL MSM jobs");
  const MsmOptions o = parse_options(options, options_len);
  ZP_TRY(validate_options(o, "zp_mixed_batch_create"));
  ZP_CHECK(!o.host_finalize, ZP_ERR_INVALID_ARG, "zp_mixed_batch_create: finalize=host cannot be captured");
  std::vector<HostJob> hj(msm_count);
  for (size_t j = 0; j < msm_count; j++) {
    const zp_msm_job &J = msm_jobs[j];
    ZP_CHECK(J.ctx != nullptr && J.out != nullptr, ZP_ERR_INVALID_ARG, "MSM job %zu: NULL context or output", j);
    ZP_CHECK(J.n <= J.ctx->n, ZP_ERR_TOO_LARGE, "MSM job %zu: n %zu > %zu points", j, J.n, J.ctx->n);
    ZP_CHECK(J.d_scalars != nullptr || J.n == 0, ZP_ERR_INVALID_ARG, "MSM job %zu: NULL scalars", j);
    ZP_CHECK(J.window == 0 || (J.window >= 8 && J.window <= 24), ZP_ERR_INVALID_ARG, "MSM job %zu: window %d", j,
             J.window);
    ZP_CHECK(!J.ctx->opts.host_finalize, ZP_ERR_INVALID_ARG,
             "MSM job %zu: its table was loaded with finalize=host, which cannot be captured", j);
    hj[j] = HostJob{J.ctx, J.n, reinterpret_cast<const uint64_t *>(J.d_scalars), J.out, J.window};
  }
  zp_mixed_batch *b = new zp_mixed_batch();
  int st = ZP_OK;
  if (ntt_count) st = zp_ntt_batch_create(ntt_ptrs, ntt_log_ns, ntt_dirs, ntt_count, &b->ntt);
  if (st == ZP_OK && msm_count) {
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
      set_last_error("zp_mixed_batch_create: stream creation failed");
      st = ZP_ERR_HIP;
    } else {
      st = plan_batch(b->ws, hj.data(), msm_count, o, s, b->plan);
      if (hipStreamSynchronize(s) != hipSuccess && st == ZP_OK) {
        set_last_error("zp_mixed_batch_create: table upload failed");
        st = ZP_ERR_HIP;
      }
      (void)hipStreamDestroy(s);
    }
    if (st == ZP_OK && hipHostMalloc((void **)&b->h_out, 18 * sizeof(uint64_t) * msm_count, hipHostMallocDefault) !=
                           hipSuccess) {
      set_last_error("zp_mixed_batch_create: pinned result buffer");
      st = ZP_ERR_OUT_OF_MEMORY;
    }
  }
  if (st == ZP_OK) st = mixed_capture(b);
  if (st != ZP_OK) {
    zp_mixed_batch_destroy(b);
    return st;
  }
  *out = b;
  return ZP_OK;
}

extern "C" int zp_mixed_batch_launch(zp_mixed_batch *b, void *stream) {
  ZP_CHECK(b != nullptr, ZP_ERR_INVALID_ARG, "zp_mixed_batch_launch: NULL batch");
  ZP_HIP(hipGraphLaunch(b->exec, (hipStream_t)stream));
  return ZP_OK;
}

extern "C" int zp_mixed_batch_wait(zp_mixed_batch *b, void *stream) {
  ZP_CHECK(b != nullptr, ZP_ERR_INVALID_ARG, "zp_mixed_batch_wait: NULL batch");
  ZP_HIP(hipStreamSynchronize((hipStream_t)stream));
  for (size_t j = 0; j < b->plan.jobs.size(); j++)
    memcpy(b->plan.jobs[j].out, b->h_out + 18 * j, sizeof(g1_projective_t));
  return ZP_OK;
}

extern "C" size_t zp_msm_num_points(const zp_msm_context *ctx) { return ctx ? ctx->n : 0; }
extern "C" int zp_msm_precompute_width(const zp_msm_context *ctx) { return ctx && ctx->d_shift ? ctx->shift_c : 0; }

extern "C" int zp_debug_msm_plan(const zp_msm_context *ctx, size_t n, int *shifted, int *window_bits,
                                 int *num_windows) {
  ZP_CHECK(ctx && shifted && window_bits && num_windows, ZP_ERR_INVALID_ARG, "zp_debug_msm_plan: null argument");
  ZP_CHECK(n <= ctx->n, ZP_ERR_INVALID_ARG, "zp_debug_msm_plan: n=%zu > %zu points", n, ctx->n);
  const HostLayout L = choose_layout(ctx->opts, n);
  HostLayout LS;
  const bool pre = job_uses_shifted(ctx, n, false, L, LS);
  *shifted = pre;
  *window_bits = pre ? ctx->shift_c : (ctx->opts.reference ? 13 : L.width[0]);
  *num_windows = pre ? LS.W : L.W;
  return ZP_OK;
}
extern "C" size_t zp_msm_num_nonconvertible(const zp_msm_context *ctx) { return ctx ? ctx->num_nonconv : 0; }

// ---------------------------------------------------------- generator --
static std::mutex g_gen_mu;
static std::map<int, PackedPre *> g_gen_table;  // device -> 32 x 256 comb table

static int gen_table(PackedPre **out) {
  int dev = 0;
  ZP_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_gen_mu);
  auto it = g_gen_table.find(dev);
  if (it != g_gen_table.end()) {
    *out = it->second;
    return ZP_OK;
  }
  const FConsts &K = consts();
  // Generator in twisted-Edwards affine form via the import path.
  const F r2 = f_from_u384(pow2_mod_q(784));
  const F gx = f_mul(f_from_u384(u384_from_hex(HEX_GX)), r2);
  const F gy = f_mul(f_from_u384(u384_from_hex(HEX_GY)), r2);
  const F u = f_mul(f_sub_kq(gx, K.alpha, K.q2), K.mont_s);
  const F v = f_mul(K.mont_s, gy);
  const F tx = f_mul(f_mul(K.ts, u), f_inv(v, K.one));
  const F ty = f_mul(f_sub_kq(u, K.one, K.q2), f_inv(f_add(u, K.one), K.one));
  Xyzt base;
  base.x = tx;
  base.y = ty;
  base.z = K.one;
  base.t = f_mul(tx, ty);
  std::vector<Xyzt> pts(32 * 256);
  for (int j = 0; j < 32; j++) {
    pts[j * 256] = ec_identity(K);
    pts[j * 256 + 1] = base;
    for (int m = 2; m < 256; m++) {
      Xyzt t = pts[j * 256 + m - 1];
      ec_add(t, base, K);
      pts[j * 256 + m] = t;
    }
    for (int d = 0; d < 8; d++) ec_dbl(base, K);
  }
  // Batch-normalise to affine and precompute.
  std::vector<F> pref(pts.size());
  F acc = K.one;
  for (size_t i = 0; i < pts.size(); i++) {
    acc = f_mul(acc, pts[i].z);
    pref[i] = acc;
  }
  F inv = f_inv(acc, K.one);
  std::vector<PackedPre> table(pts.size());
  for (size_t i = pts.size(); i-- > 0;) {
    const F zi = f_mul(inv, i ? pref[i - 1] : K.one);
    inv = f_mul(inv, pts[i].z);
    const F x = f_mul(pts[i].x, zi), y = f_mul(pts[i].y, zi);
    table[i] = pre_pack(pre_from_affine(x, y, K));
  }
  PackedPre *d = nullptr;
  ZP_HIP(hipMalloc(&d, table.size() * sizeof(PackedPre)));
  ZP_HIP(hipMemcpy(d, table.data(), table.size() * sizeof(PackedPre), hipMemcpyHostToDevice));
  g_gen_table[dev] = d;
  *out = d;
  return ZP_OK;
}

extern "C" int zp_g1_mul_generator_device(const biginteger256_t *d_k, g1_affine_t *d_out, size_t n, void *stream) {
  ZP_CHECK((d_k && d_out) || n == 0, ZP_ERR_INVALID_ARG, "zp_g1_mul_generator_device: NULL argument");
  if (n == 0) return ZP_OK;
  PackedPre *table = nullptr;
  ZP_TRY(gen_table(&table));
  const FConsts &K = consts();
  hipStream_t s = (hipStream_t)stream;
  const size_t SLICE = std::min<size_t>(n, (size_t)1 << 21);
  Xyzt *d_acc = nullptr;
  F *d_den = nullptr, *d_tmp = nullptr;
  uint8_t *d_flag = nullptr;
  int st = ZP_OK;
  if (hipMalloc(&d_acc, SLICE * sizeof(Xyzt)) != hipSuccess || hipMalloc(&d_den, SLICE * sizeof(F)) != hipSuccess ||
      hipMalloc(&d_tmp, SLICE * sizeof(F)) != hipSuccess || hipMalloc(&d_flag, SLICE) != hipSuccess) {
    set_last_error("zp_g1_mul_generator_device: allocation failed");
    st = ZP_ERR_OUT_OF_MEMORY;
  }
  for (size_t b = 0; st == ZP_OK && b < n; b += SLICE) {
    const size_t m = std::min(SLICE, n - b);
    hipLaunchKernelGGL(k_gen_a, dim3(grid_for(m)), dim3(MSM_THREADS), 0, s, d_k + b, (uint64_t)m, table, K, d_acc,
                       d_den, d_flag);
    const uint32_t chunk = 32;
    hipLaunchKernelGGL(k_batch_inv, dim3(grid_for((m + chunk - 1) / chunk, 64)), dim3(64), 0, s, d_den, (uint64_t)m,
                       chunk, d_tmp, K);
    hipLaunchKernelGGL(k_gen_c, dim3(grid_for(m)), dim3(MSM_THREADS), 0, s, d_acc, d_den, d_flag, (uint64_t)m, K,
                       d_out + b);
    if (hipGetLastError() != hipSuccess) {
      set_last_error("zp_g1_mul_generator_device: launch failed");
      st = ZP_ERR_HIP;
    }
  }
  if (hipStreamSynchronize(s) != hipSuccess && st == ZP_OK) {
    set_last_error("zp_g1_mul_generator_device: execution failed");
    st = ZP_ERR_HIP;
  }
  if (d_acc) (void)hipFree(d_acc);
  if (d_den) (void)hipFree(d_den);
  if (d_tmp) (void)hipFree(d_tmp);
  if (d_flag) (void)hipFree(d_flag);
  return st;
}

// -------------------------------------------------------- diagnostics --
extern "C" int zp_debug_fq_mul(const uint64_t *a, const uint64_t *b, uint64_t *out, size_t n) {
  ZP_CHECK((a && b && out) || n == 0, ZP_ERR_INVALID_ARG, "zp_debug_fq_mul: NULL argument");
  if (!n) return ZP_OK;
  uint64_t *d = nullptr;
  ZP_HIP(hipMalloc(&d, 3 * n * 48));
  ZP_HIP(hipMemcpy(d, a, n * 48, hipMemcpyHostToDevice));
  ZP_HIP(hipMemcpy(d + 6 * n, b, n * 48, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_debug_mul, dim3(grid_for(n)), dim3(MSM_THREADS), 0, 0, d, d + 6 * n, d + 12 * n, (uint64_t)n,
                     consts());
  ZP_HIP(hipGetLastError());
  ZP_HIP(hipMemcpy(out, d + 12 * n, n * 48, hipMemcpyDeviceToHost));
  ZP_HIP(hipFree(d));
  return ZP_OK;
}

extern "C" int zp_debug_msm_digits(const char *options, size_t options_len, const biginteger256_t *scalars, size_t n,
                                   int32_t *digits, int *num_windows) {
  ZP_CHECK(num_windows != nullptr, ZP_ERR_INVALID_ARG, "zp_debug_msm_digits: NULL num_windows");
  const MsmOptions opts = parse_options(options, options_len);
  ZP_TRY(validate_options(opts, "zp_debug_msm_digits"));
  int W = 0;
  ZP_TRY(dump_digits(opts, nullptr, n, nullptr, 0, &W));
  *num_windows = W;
  if (!digits || !n) return ZP_OK;
  ZP_CHECK(scalars != nullptr, ZP_ERR_INVALID_ARG, "zp_debug_msm_digits: NULL scalars");
  uint64_t *ds = nullptr;
  int32_t *dd = nullptr;
  ZP_HIP(hipMalloc(&ds, n * 32));
  ZP_HIP(hipMalloc(&dd, n * W * sizeof(int32_t)));
  ZP_HIP(hipMemcpy(ds, scalars, n * 32, hipMemcpyHostToDevice));
  ZP_TRY(dump_digits(opts, ds, n, dd, 0, &W));
  ZP_HIP(hipMemcpy(digits, dd, n * W * sizeof(int32_t), hipMemcpyDeviceToHost));
  ZP_HIP(hipFree(ds));
  ZP_HIP(hipFree(dd));
  return ZP_OK;
}

extern "C" int zp_debug_msm_precomputed(zp_msm_context *ctx, size_t i0, size_t count, uint64_t *out) {
  ZP_CHECK(ctx != nullptr && (out != nullptr || count == 0), ZP_ERR_INVALID_ARG, "zp_debug_msm_precomputed: NULL");
  ZP_CHECK(i0 + count <= ctx->n, ZP_ERR_INVALID_ARG, "zp_debug_msm_precomputed: range beyond %zu points", ctx->n);
  if (!count) return ZP_OK;
  uint64_t *d = nullptr;
  ZP_HIP(hipMalloc(&d, count * 18 * 8));
  hipLaunchKernelGGL(k_debug_pre_wire, dim3(grid_for(count)), dim3(MSM_THREADS), 0, 0, ctx->d_pts, (uint64_t)i0,
                     (uint64_t)count, consts(), d);
  ZP_HIP(hipGetLastError());
  ZP_HIP(hipMemcpy(out, d, count * 18 * 8, hipMemcpyDeviceToHost));
  ZP_HIP(hipFree(d));
  return ZP_OK;
}
