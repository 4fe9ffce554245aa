// Host program over the C++ NttGpuDriver mirror, with the modes of the
// reference's zprize/ntt/host programs:
//   zp_ntt_host test     (test_random.cpp)    every case against a CPU transform
//                        (test_random.cpp:56,73-90), plus the round trip; the
//                        first case is all ones when T == 1 (test_random.cpp:44-52);
//                        for --core-type REVERSE the element formula of
//                        reference.cpp:60-84.
//   zp_ntt_host bench    (bench_latency.cpp)  latency percentiles through the
//                        buffer API (bench_latency.cpp:101-117), every run's
//                        output compared with the CPU transform of the input.
//   zp_ntt_host throughput (bench_throughput.cpp) every buffer loaded once,
//                        then E evaluations enqueued round-robin over the
//                        buffers without waiting, then E waits; reports
//                        NTTs/s and checks every buffer's output
//                        (bench_throughput.cpp:60-107).
//   zp_ntt_host evaluate (evaluate_given.cpp) hex text in, "0x%016x" lines out
//                        (evaluate_given.cpp:19-76); --expected-filename adds a
//                        diff against a file in the same format.
// Flags: --log-n K | --log-row-size R (n = 4^R, the reference's form),
// --core-type NTT|REVERSE, --memory-layout NORMAL_LAYOUT|OPTIMIZED_LAYOUT and
// --log-blocks B (the FPGA's device layout and core count: accepted, results
// identical), --xclbin (accepted, unused), --num-test-cases T, --num-evaluations E, --input-filename,
// --output-filename, --expected-filename, --inverse.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/zpgpu_ntt_driver.hpp"

static const uint64_t P = 0xFFFFFFFF00000001ull;
typedef std::vector<uint64_t> vec64;

static uint64_t mulmod(uint64_t a, uint64_t b) { return (uint64_t)((unsigned __int128)a * b % P); }
static uint64_t powmod(uint64_t a, uint64_t e) {
  uint64_t r = 1;
  for (; e; e >>= 1, a = mulmod(a, a))
    if (e & 1) r = mulmod(r, a);
  return r;
}

// CPU transform the test and bench modes check the GPU against, as the
// reference's programs do with ntt_reference (test_random.cpp:56,
// bench_latency.cpp:95-98): X_k = sum_j x_j w^(jk) with w = 7^((p-1)/n)
// (w^-1 and a final n^-1 for the inverse), by bit reversal and radix-2
// decimation-in-time passes with 128-bit modular products.
static vec64 cpu_ntt(const vec64 &x, bool inverse) {
  const size_t n = x.size();
  int lg = 0;
  while ((size_t(1) << lg) < n) lg++;
  vec64 a(x);
  for (auto &v : a) v %= P;
  for (size_t i = 1, j = 0; i < n; i++) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (size_t len = 2; len <= n; len <<= 1) {
    uint64_t w = powmod(7, (P - 1) / len);
    if (inverse) w = powmod(w, P - 2);
    for (size_t i = 0; i < n; i += len) {
      uint64_t t = 1;
      for (size_t k = 0; k < len / 2; k++, t = mulmod(t, w)) {
        const uint64_t u = a[i + k], v = mulmod(a[i + k + len / 2], t);
        a[i + k] = (uint64_t)(((unsigned __int128)u + v) % P);
        a[i + k + len / 2] = (uint64_t)(((unsigned __int128)u + P - v) % P);
      }
    }
  }
  if (inverse) {
    const uint64_t ninv = powmod(n % P, P - 2);
    for (auto &v : a) v = mulmod(v, ninv);
  }
  return a;
}

// test_random.cpp:73-80: the first mismatching element.
static bool same_or_report(const vec64 &want, const vec64 &got) {
  for (size_t i = 0; i < want.size(); i++)
    if (want[i] != got[i]) {
      std::cout << "Error: Result mismatch" << std::endl
                << " i = " << i << " Expected = " << want[i] << " Obtained = " << got[i] << std::endl;
      return false;
    }
  return true;
}

// evaluate_given.cpp:19-56: exactly n hex values (an 0x prefix is accepted).
static void load_input(vec64 &data, const std::string &filename) {
  std::ifstream fin(filename);
  if (!fin.is_open()) throw std::runtime_error("Failed to open input_file \"" + filename + "\" Aborting!");
  uint64_t i = 0, x;
  while (fin >> std::hex >> x) {
    if (i >= data.size())
      throw std::runtime_error("Expecting exactly " + std::to_string(data.size()) + " elements, but got more?");
    data[i++] = x;
  }
  if (i != data.size())
    throw std::runtime_error("Expecting exactly " + std::to_string(data.size()) + " elements, in input file " +
                             filename + ", but only got " + std::to_string(i) + "?");
}

// evaluate_given.cpp:58-76
static void store_output(const vec64 &data, const std::string &filename) {
  FILE *f = fopen(filename.c_str(), "w");
  if (!f) throw std::runtime_error("Failed to open output_file \"" + filename + "\" Aborting!");
  for (uint64_t x : data) fprintf(f, "0x%016llx\n", (unsigned long long)x);
  fclose(f);
}

struct Args {
  std::string mode;
  unsigned log_n = 12;
  bool reverse = false;
  bool inverse = false;
  uint64_t cases = 2, evals = 50;
  std::string input, output, expected;
};

// Command-line errors (exit status 2, like an unknown flag in test_random.cpp).
struct UsageError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

static Args parse(int argc, char **argv) {
  Args a;
  a.mode = argv[1];
  bool have_row = false;
  for (int i = 2; i < argc; i++) {
    std::string f = argv[i];
    if (f == "--inverse") {
      a.inverse = true;
      continue;
    }
    if (i + 1 >= argc || strlen(argv[i + 1]) == 0) throw UsageError(f + " expects a non-empty argument!");
    std::string v = argv[++i];
    if (f == "--log-n") {
      a.log_n = (unsigned)std::stoul(v);
    } else if (f == "--log-row-size") {
      const unsigned r = (unsigned)std::stoul(v);
      if (r == 0) throw UsageError("--log-row-size expects a positive numerical argument!");
      a.log_n = 2 * r;
      have_row = true;
    } else if (f == "--core-type") {
      if (v == "REVERSE")
        a.reverse = true;
      else if (v != "NTT")
        throw UsageError("Unknown core type " + v);
    } else if (f == "--memory-layout") {
      // the FPGA driver's device-buffer layout; vectors are natural order in both
      if (v != "NORMAL_LAYOUT" && v != "OPTIMIZED_LAYOUT") throw UsageError("Unknown memory layout " + v);
    } else if (f == "--log-blocks") {
      (void)std::stoull(v);  // the FPGA's core count; no GPU meaning
    } else if (f == "--xclbin") {
      // the device is ready without a binary; accepted for command-line compatibility
    } else if (f == "--num-test-cases") {
      a.cases = std::stoull(v);
    } else if (f == "--num-evaluations") {
      a.evals = std::stoull(v);
    } else if (f == "--input-filename") {
      a.input = v;
    } else if (f == "--output-filename") {
      a.output = v;
    } else if (f == "--expected-filename") {
      a.expected = v;
    } else {
      throw UsageError("Unknown flag: " + f);  // test_random.cpp:189-193
    }
  }
  if (a.reverse && !have_row)
    throw UsageError("--log-row-size must be specified as a non-zero value when core_type is REVERSE");
  return a;
}

// reference.cpp:60-84 composed: out[i][j] = in[R-1-j][R-1-i] * w^(j (R-1-i)).
static bool check_reverse(const vec64 &in, const vec64 &out, unsigned log_n) {
  const unsigned log_r = log_n / 2;
  const uint64_t r = 1ull << log_r;
  const uint64_t w = powmod(7, (P - 1) >> log_n);
  for (uint64_t i = 0; i < r; i++)
    for (uint64_t j = 0; j < r; j++) {
      const uint64_t want = mulmod(in[((r - 1 - j) << log_r) + (r - 1 - i)], powmod(w, j * (r - 1 - i)));
      if (out[(i << log_r) + j] != want) {
        std::cout << "Error: Result mismatch\n i = " << ((i << log_r) + j) << " Expected = " << want
                  << " Obtained = " << out[(i << log_r) + j] << std::endl;
        return false;
      }
    }
  return true;
}

static int run(const Args &a) {
  const zpgpu::NttGpuDriverArg arg =
      a.reverse ? zpgpu::NttGpuDriverArg::create_reverse(a.log_n / 2) : zpgpu::NttGpuDriverArg::create(a.log_n);
  zpgpu::NttGpuDriver drv(arg);
  const uint64_t n = drv.input_vector_size();
  vec64 in(n), fwd(n), back(n);
  std::mt19937_64 rng;
  const int dir = a.inverse ? ZP_NTT_INVERSE : ZP_NTT_FORWARD;

  if (a.mode == "evaluate") {
    if (a.input.empty() || a.output.empty()) throw std::runtime_error("evaluate needs --input-filename and --output-filename");
    load_input(in, a.input);
    drv.simple_evaluate(fwd.data(), in.data(), n, dir);
    store_output(fwd, a.output);
    if (a.expected.empty()) return 0;
    vec64 want(n);
    load_input(want, a.expected);
    for (uint64_t i = 0; i < n; i++)
      if (want[i] != fwd[i]) {
        std::cout << "Error: Result mismatch\n i = " << i << " Expected = " << want[i] << " Obtained = " << fwd[i]
                  << "\nEVALUATE FAILED" << std::endl;
        return 1;
      }
    std::cout << "EVALUATE PASSED" << std::endl;
    return 0;
  }

  if (a.mode == "test") {
    int bad = 0;
    for (uint64_t t = 0; t < a.cases; t++) {
      for (uint64_t i = 0; i < n; i++) {
        if (a.cases == 1 && t == 0) {
          in[i] = 1;
        } else {
          uint64_t v;
          do v = rng(); while (v >= P);
          in[i] = v;
        }
      }
      bool ok;
      if (a.reverse) {
        drv.simple_evaluate(fwd.data(), in.data(), n, ZP_NTT_FORWARD);
        ok = check_reverse(in, fwd, a.log_n);
      } else {
        drv.simple_evaluate(fwd.data(), in.data(), n, ZP_NTT_FORWARD);
        drv.simple_evaluate(back.data(), fwd.data(), n, ZP_NTT_INVERSE);
        ok = same_or_report(cpu_ntt(in, false), fwd) && back == in;
        if (a.cases == 1 && t == 0) {  // all ones -> n * delta
          ok = ok && fwd[0] == n % P;
          for (uint64_t i = 1; i < n; i++) ok = ok && fwd[i] == 0;
        }
      }
      printf("test case %llu: %s\n", (unsigned long long)t, ok ? "OK" : "FAILED");
      bad += !ok;
    }
    printf("\nNTT TEST %s\n", bad ? "FAILED" : "PASSED");
    return bad ? 1 : 0;
  }

  if (a.mode == "bench") {
    // request / transfer / enqueue / wait / transfer / free, per evaluation
    // one input, its CPU transform computed once, every evaluation compared
    // with it outside the timed region (bench_latency.cpp:60-98)
    std::vector<double> lat;
    for (uint64_t i = 0; i < n; i++) in[i] = rng() % P;
    std::cout << "Computing reference output.." << std::endl;
    const vec64 expected = cpu_ntt(in, dir == ZP_NTT_INVERSE);
    int failed = 0;
    for (uint64_t e = 0; e < a.evals; e++) {
      auto t0 = std::chrono::steady_clock::now();
      auto *b = drv.request_buffer();
      drv.transfer_input_vector_to_user_buffer(b, in.data());
      drv.enqueue_evaluation_async(b, dir);
      drv.wait_for_result(b);
      drv.transfer_user_buffer_to_output_vector(b, fwd.data());
      drv.free_buffer(b);
      lat.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
      if (fwd != expected) {
        std::cout << "Incorrect result in run  " << e << std::endl;
        failed = 1;
      }
    }
    std::sort(lat.begin(), lat.end());
    auto q = [&](double f) { return lat[(size_t)(f * (lat.size() - 1))]; };
    printf("latency (s): min %.6f p25 %.6f p50 %.6f p75 %.6f max %.6f\n", lat.front(), q(0.25), q(0.5), q(0.75),
           lat.back());
    return failed;
  }
  if (a.mode == "throughput") {
    const uint64_t nb = drv.max_num_buffers_in_flight();
    std::vector<vec64> inputs(nb, vec64(n)), want(nb, vec64(n));
    std::vector<zpgpu::NttGpuDriver::UserBuffer *> bufs;
    for (uint64_t t = 0; t < nb; t++) {
      for (uint64_t i = 0; i < n; i++) {
        uint64_t v;
        do v = rng(); while (v >= P);
        inputs[t][i] = v;
      }
      bufs.push_back(drv.request_buffer());
      drv.transfer_input_vector_to_user_buffer(bufs[t], inputs[t].data());
      // expected output: one blocking evaluation of the same input
      drv.enqueue_evaluation_async(bufs[t], dir);
      drv.wait_for_result(bufs[t]);
      drv.transfer_user_buffer_to_output_vector(bufs[t], want[t].data());
    }
    const auto t0 = std::chrono::steady_clock::now();
    for (uint64_t e = 0, t = 0; e < a.evals; e++, t = (t + 1) % nb) drv.enqueue_evaluation_async(bufs[t], dir);
    for (uint64_t e = 0, t = 0; e < a.evals; e++, t = (t + 1) % nb) drv.wait_for_result(bufs[t]);
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    int bad = 0;
    for (uint64_t t = 0; t < nb; t++) {
      drv.transfer_user_buffer_to_output_vector(bufs[t], fwd.data());
      if (fwd != want[t]) {
        std::cout << "Incorrect result in buffer " << t << std::endl;
        bad = 1;
      }
      drv.free_buffer(bufs[t]);
    }
    printf("Time taken for %llu NTTs: %.6fs\nAmortized time taken per NTT %.6fs\nThroughput: %.3f NTTs/second\n",
           (unsigned long long)a.evals, secs, secs / (double)a.evals, (double)a.evals / secs);
    return bad;
  }
  throw std::runtime_error("unknown mode " + a.mode);
}

int main(int argc, char **argv) {
  if (argc < 2) {
    fprintf(stderr,
            "usage: %s test|bench|throughput|evaluate [--log-n K | --log-row-size R] [--core-type NTT|REVERSE]\n"
            "       [--num-test-cases T] [--num-evaluations E] [--input-filename F --output-filename F\n"
            "       [--expected-filename F]] [--inverse]\n",
            argv[0]);
    return 2;
  }
  try {
    return run(parse(argc, argv));
  } catch (const UsageError &ex) {
    fprintf(stderr, "error: %s\n", ex.what());
    return 2;
  } catch (const std::exception &ex) {
    fprintf(stderr, "error: %s\n", ex.what());
    return 1;
  }
}
