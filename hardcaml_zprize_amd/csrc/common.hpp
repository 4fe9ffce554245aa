// Shared host-side helpers: status codes, HIP error plumbing, thread-local
// last-error text for the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/zpgpu.h"

namespace zp {

void set_last_error(const char *fmt, ...);

#define ZP_HIP(expr)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) {                                                             \
      ::zp::set_last_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                           __LINE__);                                                   \
      return ZP_ERR_HIP;                                                                \
    }                                                                                   \
  } while (0)

#define ZP_CHECK(cond, code, ...)       \
  do {                                  \
    if (!(cond)) {                      \
      ::zp::set_last_error(__VA_ARGS__); \
      return (code);                    \
    }                                   \
  } while (0)

#define ZP_TRY(expr)            \
  do {                          \
    int s_ = (expr);            \
    if (s_ != ZP_OK) return s_; \
  } while (0)

}  // namespace zp
