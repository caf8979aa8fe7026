// Shared host-side helpers of librcv: thread-local error reporting, HIP
// status checking and launch-shape validation.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/rcv.h"

namespace rcv {

// Thread-local message of the last failing C-ABI call on this thread.
std::string &last_error_slot();

inline int fail(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  last_error_slot() = buf;
  return code;
}

inline int hip_fail(hipError_t e, const char *what) {
  return fail(RCV_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Checks the launch just enqueued (configuration errors surface here).
inline int check_launch(const char *name) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, name);
  return RCV_OK;
}

#define RCV_TRY_HIP(expr, what)                 \
  do {                                          \
    hipError_t _e = (expr);                     \
    if (_e != hipSuccess) return ::rcv::hip_fail(_e, what); \
  } while (0)

#define RCV_REQUIRE(cond, ...)                                \
  do {                                                        \
    if (!(cond)) return ::rcv::fail(RCV_E_INVALID, __VA_ARGS__); \
  } while (0)

inline bool aligned16(const void *p) {
  return (reinterpret_cast<uintptr_t>(p) & 15u) == 0;
}

inline unsigned ceil_div(int64_t a, int64_t b) { return unsigned((a + b - 1) / b); }

}  // namespace rcv
