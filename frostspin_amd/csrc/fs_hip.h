// HIP-side helpers shared by the kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "fs_common.h"

#define FS_HIP_CHECK(expr)                                                     \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess) {                                                    \
      fs::set_error("%s failed: %s", #expr, hipGetErrorString(e_));           \
      return FS_ERR_HIP;                                                       \
    }                                                                          \
  } while (0)

#define FS_LAUNCH_CHECK(name)                                                  \
  do {                                                                         \
    hipError_t e_ = hipGetLastError();                                         \
    if (e_ != hipSuccess) {                                                    \
      fs::set_error("launch of %s failed: %s", name, hipGetErrorString(e_));  \
      return FS_ERR_HIP;                                                       \
    }                                                                          \
  } while (0)

namespace fs {
inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }
inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
} // namespace fs
