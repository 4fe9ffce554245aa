// Host program over the C++ NttGpuDriver mirror, with the modes of the
// reference's zprize/ntt/host/test_random.cpp and bench_latency.cpp:
//   zp_ntt_host test  --log-n K --num-test-cases T   (round trip + closed forms;
//                     the first case is all ones when T == 1, test_random.cpp:44-52)
//   zp_ntt_host bench --log-n K --num-evaluations E  (latency percentiles through
//                     the buffer API, bench_latency.cpp:101-117)
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "../../../include/zpgpu_ntt_driver.hpp"

static const uint64_t P = 0xFFFFFFFF00000001ull;

int main(int argc, char **argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: %s test|bench [--log-n K] [--num-test-cases T] [--num-evaluations E]\n", argv[0]);
    return 2;
  }
  std::string mode = argv[1];
  unsigned log_n = 12;
  uint64_t cases = 2, evals = 50;
  for (int i = 2; i + 1 < argc; i += 2) {
    if (!strcmp(argv[i], "--log-n"))
      log_n = (unsigned)atoi(argv[i + 1]);
    else if (!strcmp(argv[i], "--num-test-cases"))
      cases = strtoull(argv[i + 1], nullptr, 10);
    else if (!strcmp(argv[i], "--num-evaluations"))
      evals = strtoull(argv[i + 1], nullptr, 10);
    else {
      fprintf(stderr, "unknown flag %s\n", argv[i]);  // test_random.cpp:189-193
      return 2;
    }
  }
  try {
    zpgpu::NttGpuDriver drv(zpgpu::NttGpuDriverArg::create(log_n));
    const uint64_t n = drv.input_vector_size();
    std::vector<uint64_t> in(n), fwd(n), back(n);
    std::mt19937_64 rng;
    if (mode == "test") {
      int bad = 0;
      for (uint64_t t = 0; t < cases; t++) {
        for (uint64_t i = 0; i < n; i++) {
          if (cases == 1 && t == 0) {
            in[i] = 1;
          } else {
            uint64_t v;
            do v = rng(); while (v >= P);
            in[i] = v;
          }
        }
        drv.simple_evaluate(fwd.data(), in.data(), n, ZP_NTT_FORWARD);
        drv.simple_evaluate(back.data(), fwd.data(), n, ZP_NTT_INVERSE);
        bool ok = back == in;
        if (cases == 1 && t == 0) {  // all ones -> n * delta
          ok = ok && fwd[0] == n % P;
          for (uint64_t i = 1; i < n; i++) ok = ok && fwd[i] == 0;
        }
        printf("test case %llu: %s\n", (unsigned long long)t, ok ? "OK" : "FAILED");
        bad += !ok;
      }
      return bad ? 1 : 0;
    }
    // bench: request / transfer / enqueue / wait / transfer / free, per evaluation
    std::vector<double> lat;
    for (uint64_t e = 0; e < evals; e++) {
      for (uint64_t i = 0; i < n; i++) in[i] = rng() % P;
      auto t0 = std::chrono::steady_clock::now();
      auto *b = drv.request_buffer();
      drv.transfer_input_vector_to_user_buffer(b, in.data());
      drv.enqueue_evaluation_async(b);
      drv.wait_for_result(b);
      drv.transfer_user_buffer_to_output_vector(b, fwd.data());
      drv.free_buffer(b);
      lat.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    }
    std::sort(lat.begin(), lat.end());
    auto q = [&](double f) { return lat[(size_t)(f * (lat.size() - 1))]; };
    printf("latency (s): min %.6f p25 %.6f p50 %.6f p75 %.6f max %.6f\n", lat.front(), q(0.25), q(0.5), q(0.75),
           lat.back());
  } catch (const std::exception &ex) {
    fprintf(stderr, "error: %s\n", ex.what());
    return 1;
  }
  return 0;
}
