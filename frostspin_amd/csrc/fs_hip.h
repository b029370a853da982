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

#include <mutex>

namespace fs {
inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }
inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline size_t align_up(size_t a, size_t b) { return (a + b - 1) / b * b; }

// Device copy of a plan's read-only tables. Neither the upload nor the release
// synchronises a compute stream: the memory comes from the device's
// stream-ordered pool through a per-device upload stream that is
// non-blocking (it does not wait for the legacy default stream), the upload
// waits for its own copy only, and release() frees the memory in stream order
// after the work already queued on every stream that executed the plan
// (note_stream() records them; beyond kMaxStreams the release waits for the
// device instead). A synchronous hipMemcpy on the default stream would drain
// every kernel queued there, and hipFree waits for the whole device.
class DeviceTables {
 public:
  static constexpr int kMaxStreams = 8;
  DeviceTables() = default;
  DeviceTables(const DeviceTables &) = delete;
  DeviceTables &operator=(const DeviceTables &) = delete;
  ~DeviceTables() { (void)release(); }

  int upload(const void *host, size_t bytes);  // on the current device
  void note_stream(hipStream_t s);             // before every launch that reads the tables
  int release();                               // idempotent
  const char *data() const { return static_cast<const char *>(ptr_); }
  int device() const { return device_; }

 private:
  void *ptr_ = nullptr;
  int device_ = -1;
  std::mutex mu_;
  hipStream_t streams_[kMaxStreams] = {};
  int nstreams_ = 0;
  bool overflow_ = false;
};
} // namespace fs
