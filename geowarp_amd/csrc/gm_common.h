// Shared host/device helpers of libgeowarp_amd: error reporting and launch utilities.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

namespace gm {
void set_error(const char* fmt, ...);
}

#define GM_HIP(call)                                                                  \
  do {                                                                               \
    hipError_t _e = (call);                                                          \
    if (_e != hipSuccess) {                                                          \
      gm::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(_e)); \
      return -2;                                                                     \
    }                                                                                \
  } while (0)

#define GM_LAUNCH_CHECK()                                                                \
  do {                                                                                  \
    hipError_t _e = hipGetLastError();                                                  \
    if (_e != hipSuccess) {                                                             \
      gm::set_error("%s:%d kernel launch: %s", __FILE__, __LINE__, hipGetErrorString(_e)); \
      return -3;                                                                        \
    }                                                                                   \
  } while (0)

#define GM_REQUIRE(cond, msg)            \
  do {                                   \
    if (!(cond)) {                       \
      gm::set_error("%s", msg);          \
      return -1;                         \
    }                                    \
  } while (0)

static inline int gm_blocks(long n, int t) { return (int)((n + t - 1) / t); }
