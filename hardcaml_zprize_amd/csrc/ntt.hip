// Goldilocks NTT on gfx950: plans, kernels and the NTT half of the C ABI.
//
// A transform of n = 2^k points:
//   k <= 12: one workgroup holds the whole vector in LDS (a single pass);
//   k  > 12: the 4-step of libs/hardcaml_ntt/src/reference_model.ml:92-105 with
//            n = n1 * n2 (n1 = 2^ceil(k/2)), as the FPGA's two phases
//            (zprize/ntt/host/driver.cpp:201-231):
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
  std::vector<uint64_t> sub(2048), lo(4096), hi(n_hi);
  for (int dir = 0; dir < 2; dir++) {
    uint64_t w4096 = gl_root_host(12);
    uint64_t wn = gl_root_host(log_n);
    if (dir) {
      w4096 = gl_inv_host(w4096);
      wn = gl_inv_host(wn);
    }
    uint64_t acc = 1;
    for (int i = 0; i < 2048; i++, acc = gl_mul_host(acc, w4096)) sub[i] = acc;
    acc = 1;
    for (int i = 0; i < 4096; i++, acc = gl_mul_host(acc, wn)) lo[i] = acc;
    uint64_t step = gl_pow_host(wn, 4096);
    acc = 1;
    for (uint64_t i = 0; i < n_hi; i++, acc = gl_mul_host(acc, step)) hi[i] = acc;
    ZP_HIP(hipMalloc(&t->d_sub[dir], 2048 * 8));
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
};

static NttShape shape_of(unsigned log_n) {
  NttShape s;
  s.log_n = log_n;
  if (log_n <= 12) {
    s.log_n1 = log_n;
    s.log_n2 = 0;
  } else {
    s.log_n1 = (log_n + 1) / 2;
    s.log_n2 = log_n - s.log_n1;
  }
  return s;
}

// Enqueue one in-place transform.  scratch must hold n elements when log_n > 12.
static int launch_transform(const NttTables *t, uint64_t *d_data, uint64_t *d_scratch, int dir,
                            hipStream_t stream) {
  const NttShape sh = shape_of(t->log_n);
  const uint64_t n = 1ull << sh.log_n;
  const uint64_t scale = dir ? t->n_inv : 1;
  if (sh.log_n2 == 0) {
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
    a.scale = scale;
    return launch_dft(a, 1, stream);
  }
  const uint64_t n1 = 1ull << sh.log_n1, n2 = 1ull << sh.log_n2;
  // Pass 1: columns (length n1, n2 of them) + twiddle, data -> scratch.
  DftArgs p1{};
  p1.in = d_data;
  p1.out = d_scratch;
  p1.log_L = sh.log_n1;
  p1.log_T = dft_log_T(sh.log_n1, sh.log_n2);
  p1.in_elem_stride = n2;
  p1.in_seq_stride = 1;
  p1.out_elem_stride = n2;
  p1.out_seq_stride = 1;
  p1.tw_sub = t->d_sub[dir];
  p1.tw_lo = t->d_lo[dir];
  p1.tw_hi = t->d_hi[dir];
  p1.scale = 1;
  ZP_TRY(launch_dft(p1, n2 >> p1.log_T, stream));
  // Pass 2: rows (length n2, n1 of them), scratch -> data transposed.
  DftArgs p2{};
  p2.in = d_scratch;
  p2.out = d_data;
  p2.log_L = sh.log_n2;
  p2.log_T = dft_log_T(sh.log_n2, sh.log_n1);
  p2.in_elem_stride = 1;
  p2.in_seq_stride = n2;
  p2.out_elem_stride = n1;
  p2.out_seq_stride = 1;
  p2.tw_sub = t->d_sub[dir];
  p2.scale = scale;
  return launch_dft(p2, n1 >> p2.log_T, stream);
}

}  // namespace zp

using namespace zp;

// ------------------------------------------------------------------ plan --
struct UserBuffer {
  uint64_t *h_in = nullptr;
  uint64_t *h_out = nullptr;
  uint64_t *d_buf = nullptr;
  hipEvent_t done = nullptr;
  bool in_use = false;
  bool enqueued = false;
};

struct zp_ntt_plan {
  unsigned log_n = 0;
  uint64_t n = 0;
  NttTables *tables = nullptr;
  uint64_t *d_scratch = nullptr;  // n elements (4-step intermediate)
  uint64_t *d_work = nullptr;     // n elements for simple_evaluate
  hipStream_t stream = nullptr;
  std::vector<UserBuffer> buffers;
};

extern "C" const char *zp_last_error(void) { return g_last_error.c_str(); }
extern "C" const char *zp_version(void) { return "zpgpu gfx950"; }

extern "C" int zp_ntt_create(unsigned log_n, unsigned num_buffers, zp_ntt_plan **out) {
  ZP_CHECK(out != nullptr, ZP_ERR_INVALID_ARG, "zp_ntt_create: out is NULL");
  *out = nullptr;
  ZP_CHECK(log_n >= 1 && log_n <= 24, ZP_ERR_INVALID_ARG, "zp_ntt_create: log_n %u not in [1, 24]", log_n);
  if (num_buffers == 0) num_buffers = 8;
  ZP_CHECK(num_buffers <= 64, ZP_ERR_INVALID_ARG, "zp_ntt_create: num_buffers %u > 64", num_buffers);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    set_last_error("zp_ntt_create: no HIP device");
    return ZP_ERR_NO_DEVICE;
  }
  zp_ntt_plan *p = new zp_ntt_plan();
  p->log_n = log_n;
  p->n = 1ull << log_n;
  int st = make_tables(log_n, &p->tables);
  if (st == ZP_OK && log_n > 12 && hipMalloc(&p->d_scratch, p->n * 8) != hipSuccess) st = ZP_ERR_OUT_OF_MEMORY;
  if (st == ZP_OK && hipMalloc(&p->d_work, p->n * 8) != hipSuccess) st = ZP_ERR_OUT_OF_MEMORY;
  if (st == ZP_OK && hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) st = ZP_ERR_HIP;
  if (st == ZP_OK) {
    p->buffers.resize(num_buffers);
  }
  if (st != ZP_OK) {
    if (st == ZP_ERR_OUT_OF_MEMORY) set_last_error("zp_ntt_create: device allocation failed");
    zp_ntt_destroy(p);
    return st;
  }
  *out = p;
  return ZP_OK;
}

extern "C" int zp_ntt_destroy(zp_ntt_plan *p) {
  if (!p) return ZP_OK;
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  for (auto &b : p->buffers) {
    if (b.h_in) (void)hipHostFree(b.h_in);
    if (b.h_out) (void)hipHostFree(b.h_out);
    if (b.d_buf) (void)hipFree(b.d_buf);
    if (b.done) (void)hipEventDestroy(b.done);
  }
  if (p->d_scratch) (void)hipFree(p->d_scratch);
  if (p->d_work) (void)hipFree(p->d_work);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
  return ZP_OK;
}

extern "C" uint64_t zp_ntt_size(const zp_ntt_plan *p) { return p ? p->n : 0; }
extern "C" unsigned zp_ntt_max_buffers_in_flight(const zp_ntt_plan *p) {
  return p ? (unsigned)p->buffers.size() : 0;
}

extern "C" int zp_ntt_inplace_device(zp_ntt_plan *p, uint64_t *d_data, int dir, void *stream) {
  ZP_CHECK(p != nullptr && d_data != nullptr, ZP_ERR_INVALID_ARG, "zp_ntt_inplace_device: NULL argument");
  ZP_CHECK(dir == 0 || dir == 1, ZP_ERR_INVALID_ARG, "zp_ntt_inplace_device: dir %d not 0/1", dir);
  return launch_transform(p->tables, d_data, p->d_scratch, dir, (hipStream_t)stream);
}

extern "C" int zp_ntt_simple_evaluate(zp_ntt_plan *p, uint64_t *out, const uint64_t *in, uint64_t len, int dir) {
  ZP_CHECK(p != nullptr && out != nullptr && (in != nullptr || len == 0), ZP_ERR_INVALID_ARG,
           "zp_ntt_simple_evaluate: NULL argument");
  ZP_CHECK(dir == 0 || dir == 1, ZP_ERR_INVALID_ARG, "zp_ntt_simple_evaluate: dir %d not 0/1", dir);
  ZP_CHECK(len <= p->n, ZP_ERR_TOO_LARGE, "zp_ntt_simple_evaluate: data_length %llu > %llu",
           (unsigned long long)len, (unsigned long long)p->n);
  if (len) ZP_HIP(hipMemcpyAsync(p->d_work, in, len * 8, hipMemcpyHostToDevice, p->stream));
  if (len < p->n) ZP_HIP(hipMemsetAsync(p->d_work + len, 0, (p->n - len) * 8, p->stream));
  ZP_TRY(launch_transform(p->tables, p->d_work, p->d_scratch, dir, p->stream));
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

extern "C" int zp_ntt_enqueue(zp_ntt_plan *p, int id, int dir) {
  ZP_TRY(check_buffer(p, id));
  UserBuffer &b = p->buffers[id];
  ZP_CHECK(b.in_use, ZP_ERR_STATE, "zp_ntt_enqueue: buffer %d not requested", id);
  ZP_CHECK(!b.enqueued, ZP_ERR_STATE, "zp_ntt_enqueue: buffer %d already enqueued", id);
  ZP_CHECK(dir == 0 || dir == 1, ZP_ERR_INVALID_ARG, "zp_ntt_enqueue: dir %d not 0/1", dir);
  ZP_HIP(hipMemcpyAsync(b.d_buf, b.h_in, p->n * 8, hipMemcpyHostToDevice, p->stream));
  ZP_TRY(launch_transform(p->tables, b.d_buf, p->d_scratch, dir, p->stream));
  ZP_HIP(hipMemcpyAsync(b.h_out, b.d_buf, p->n * 8, hipMemcpyDeviceToHost, p->stream));
  ZP_HIP(hipEventRecord(b.done, p->stream));
  b.enqueued = true;
  return ZP_OK;
}

extern "C" int zp_ntt_wait(zp_ntt_plan *p, int id) {
  ZP_TRY(check_buffer(p, id));
  UserBuffer &b = p->buffers[id];
  ZP_CHECK(b.in_use && b.enqueued, ZP_ERR_STATE, "zp_ntt_wait: buffer %d not enqueued", id);
  ZP_HIP(hipEventSynchronize(b.done));
  b.enqueued = false;
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
static std::mutex g_batch_mu;
static std::map<int, std::pair<uint64_t *, size_t>> g_batch_scratch;  // device -> arena

extern "C" int zp_ntt_batch_device(uint64_t *const *d_ptrs, const unsigned *log_ns, const int *dirs, size_t count,
                                   void *stream) {
  ZP_CHECK(count == 0 || (d_ptrs && log_ns && dirs), ZP_ERR_INVALID_ARG, "zp_ntt_batch_device: NULL argument");
  size_t scratch_need = 0;
  for (size_t i = 0; i < count; i++) {
    ZP_CHECK(log_ns[i] >= 1 && log_ns[i] <= 24, ZP_ERR_INVALID_ARG, "job %zu: log_n %u not in [1, 24]", i, log_ns[i]);
    ZP_CHECK(dirs[i] == 0 || dirs[i] == 1, ZP_ERR_INVALID_ARG, "job %zu: dir %d", i, dirs[i]);
    ZP_CHECK(d_ptrs[i] != nullptr, ZP_ERR_INVALID_ARG, "job %zu: NULL data", i);
    if (log_ns[i] > 12) scratch_need += 1ull << log_ns[i];
  }
  int dev = 0;
  ZP_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_batch_mu);
  auto &arena = g_batch_scratch[dev];
  if (arena.second < scratch_need) {
    if (arena.first) {
      ZP_HIP(hipStreamSynchronize((hipStream_t)stream));
      ZP_HIP(hipFree(arena.first));
      arena.first = nullptr;
      arena.second = 0;
    }
    ZP_HIP(hipMalloc(&arena.first, scratch_need * 8));
    arena.second = scratch_need;
  }
  size_t off = 0;
  for (size_t i = 0; i < count; i++) {
    NttTables *t = nullptr;
    ZP_TRY(make_tables(log_ns[i], &t));
    uint64_t *scr = nullptr;
    if (log_ns[i] > 12) {
      scr = arena.first + off;
      off += 1ull << log_ns[i];
    }
    ZP_TRY(launch_transform(t, d_ptrs[i], scr, dirs[i], (hipStream_t)stream));
  }
  return ZP_OK;
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
