// C++ host mirror of the reference's NttFpgaDriver (zprize/ntt/host/driver.h:58-225)
// over the zpgpu C ABI: same method names, argument meaning and error
// behaviour (std::runtime_error where the reference throws), so the
// reference's host programs (test_random.cpp, evaluate_given.cpp,
// bench_latency.cpp, bench_throughput.cpp) port by swapping the type.
// Header-only; link against hardcaml_zprize_amd/libzpgpu.so.
#pragma once
#include <stdint.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "zpgpu.h"

namespace zpgpu {

// driver.h:24-27.  The FPGA driver permutes vectors into OPTIMIZED_LAYOUT
// device buffers and back (driver.cpp:251-277, ntt_preprocessing.cpp,
// ntt_postprocessing.cpp); callers hand it and get back natural-order
// vectors in both layouts.  The GPU transform works on the natural layout
// directly, so both layouts behave identically here.
enum class MemoryLayout { NORMAL_LAYOUT, OPTIMIZED_LAYOUT };

// NttFpgaDriverArg::create_ntt_2_24 / _2_18 / _2_12 / create_reverse
// (driver.h:36-56), with the same arguments; log_blocks sizes the FPGA's
// core array and has no GPU meaning.  create(log_n) takes any size 2^1..2^24.
struct NttGpuDriverArg {
  unsigned log_n;
  bool reverse = false;
  MemoryLayout memory_layout = MemoryLayout::NORMAL_LAYOUT;
  uint64_t log_blocks = 0;
  static NttGpuDriverArg create_ntt_2_24(MemoryLayout l = MemoryLayout::NORMAL_LAYOUT, uint64_t log_blocks = 0) {
    return {24, false, l, log_blocks};
  }
  static NttGpuDriverArg create_ntt_2_18(MemoryLayout l = MemoryLayout::NORMAL_LAYOUT, uint64_t log_blocks = 0) {
    return {18, false, l, log_blocks};
  }
  static NttGpuDriverArg create_ntt_2_12(MemoryLayout l = MemoryLayout::NORMAL_LAYOUT, uint64_t log_blocks = 0) {
    return {12, false, l, log_blocks};
  }
  static NttGpuDriverArg create_reverse(MemoryLayout l, uint64_t log_row_size, uint64_t log_blocks = 0) {
    return {(unsigned)(2 * log_row_size), true, l, log_blocks};
  }
  static NttGpuDriverArg create_reverse(unsigned log_row_size) {
    return create_reverse(MemoryLayout::NORMAL_LAYOUT, log_row_size);
  }
  static NttGpuDriverArg create(unsigned log_n) { return {log_n}; }
  uint64_t row_size() const { return 1ull << (log_n / 2); }
  uint64_t num_elements() const { return 1ull << log_n; }
};

class NttGpuDriver {
 public:
  class UserBuffer {
   public:
    uint64_t *input_data() { return zp_ntt_buffer_input(plan_, id_); }
    uint64_t *output_data() { return zp_ntt_buffer_output(plan_, id_); }
    int id() const { return id_; }

   private:
    friend class NttGpuDriver;
    UserBuffer(zp_ntt_plan *p, int id) : plan_(p), id_(id) {}
    zp_ntt_plan *plan_;
    int id_;
  };

  explicit NttGpuDriver(NttGpuDriverArg arg, unsigned num_buffers = 8) {
    if (arg.reverse)
      check(zp_ntt_create_reverse(arg.log_n / 2, num_buffers, &plan_), "NttGpuDriver");
    else
      check(zp_ntt_create(arg.log_n, num_buffers, &plan_), "NttGpuDriver");
    for (unsigned i = 0; i < zp_ntt_max_buffers_in_flight(plan_); i++) buffers_.push_back(UserBuffer(plan_, (int)i));
  }
  ~NttGpuDriver() { zp_ntt_destroy(plan_); }
  NttGpuDriver(const NttGpuDriver &) = delete;
  NttGpuDriver &operator=(const NttGpuDriver &) = delete;

  // The device is ready at construction; kept for source compatibility
  // with driver.h:120-134 (the binary path is ignored).
  void load_xclbin(const std::string &) {}

  uint64_t input_vector_size() { return zp_ntt_size(plan_); }
  uint64_t max_num_buffers_in_flight() { return zp_ntt_max_buffers_in_flight(plan_); }

  // NULL when every buffer is in flight (driver.h:146-150).
  UserBuffer *request_buffer() {
    int id = -1;
    int st = zp_ntt_request_buffer(plan_, &id);
    if (st == ZP_ERR_NO_BUFFER) return nullptr;
    check(st, "request_buffer");
    return &buffers_[id];
  }

  void transfer_input_vector_to_user_buffer(UserBuffer *b, const uint64_t *in) {
    uint64_t *dst = b->input_data();
    for (uint64_t i = 0, n = input_vector_size(); i < n; i++) dst[i] = in[i];
  }

  void enqueue_evaluation_async(UserBuffer *b, int dir = ZP_NTT_FORWARD) {
    check(zp_ntt_enqueue(plan_, b->id_, dir), "enqueue_evaluation_async");
  }

  void wait_for_result(UserBuffer *b) { check(zp_ntt_wait(plan_, b->id_), "wait_for_result"); }

  void transfer_user_buffer_to_output_vector(UserBuffer *b, uint64_t *out) {
    const uint64_t *src = b->output_data();
    for (uint64_t i = 0, n = input_vector_size(); i < n; i++) out[i] = src[i];
  }

  // runtime_error on a double free (driver.cpp:160-170).
  void free_buffer(UserBuffer *b) { check(zp_ntt_free_buffer(plan_, b->id_), "free_buffer"); }

  // Zero-pads short input; runtime_error if data_length > n (driver.cpp:289-297, 332-344).
  void simple_evaluate(uint64_t *output, const uint64_t *input, uint64_t data_length, int dir = ZP_NTT_FORWARD) {
    check(zp_ntt_simple_evaluate(plan_, output, input, data_length, dir), "simple_evaluate");
  }

  // In-place transform of a vector already in HBM (no host copies).
  void evaluate_device(uint64_t *d_data, int dir = ZP_NTT_FORWARD, void *stream = nullptr) {
    check(zp_ntt_inplace_device(plan_, d_data, dir, stream), "evaluate_device");
  }

 private:
  static void check(int status, const char *what) {
    if (status != ZP_OK) throw std::runtime_error(std::string(what) + ": " + zp_last_error());
  }
  zp_ntt_plan *plan_ = nullptr;
  std::vector<UserBuffer> buffers_;
};

typedef NttGpuDriver::UserBuffer NttGpuBuffer;

}  // namespace zpgpu
