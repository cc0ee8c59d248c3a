// Shared helpers of the host library: status codes, the per-thread error
// message and HIP error checking.  No part of this file runs on the device.
#pragma once

#include <hip/hip_runtime.h>

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
