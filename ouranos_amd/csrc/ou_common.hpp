// Shared helpers of the host library: status codes, the per-thread error
// message and HIP error checking.  No part of this file runs on the device.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "../../include/ouranos_amd.h"

namespace ou {

// Exception carrying an ou_status; converted to a return code at the C ABI.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &msg) : std::runtime_error(msg), code(c) {}
};

void set_last_error(const std::string &msg);

[[noreturn]] inline void fail(int code, const std::string &msg) {
  throw Error(code, msg);
}

// Verbose solver progress: printed and flushed at once, so that a caller
// reading redirected output sees every line as it happens.
inline void progress(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
inline void progress(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vprintf(fmt, ap);
  va_end(ap);
  std::fflush(stdout);
}

inline void hip_check(hipError_t e, const char *what, const char *file, int line) {
  if (e != hipSuccess) {
    char buf[512];
    std::snprintf(buf, sizeof buf, "%s failed at %s:%d: %s", what, file, line,
                  hipGetErrorString(e));
    throw Error(OU_ERR_HIP, buf);
  }
}

}  // namespace ou

#define OU_HIP(call) ::ou::hip_check((call), #call, __FILE__, __LINE__)
