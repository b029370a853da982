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
#include <vector>

namespace fs {
inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }
inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline size_t align_up(size_t a, size_t b) { return (a + b - 1) / b * b; }

// Device copy of a plan's read-only tables. No call here waits for the device
// in the steady state: tables up to kArenaMaxBytes come from a per-device
// arena of size-classed blocks (device blocks plus pinned host staging blocks,
// carved from slabs that are kept until the process ends). upload() copies the
// host image into a staging block; the first prepare() queues the host-to-device
// copy on the stream that first executes the plan, so stream order puts it
// before the plan's kernels; a later prepare() on another stream makes that
// stream wait for the copy unless it is known to be complete. release() records
// one event on each stream that executed the plan and hands the blocks back to
// the arena, which reuses them once those events (and the copy's) have
// completed (beyond kMaxStreams streams the release waits for the device).
// Every stream passed to prepare() must stay valid until the tables are
// released. Larger tables use the device's stream-ordered pool on a
// non-blocking upload stream (copy waited for at upload, free ordered after
// every stream that executed the plan).
class DeviceTables {
 public:
  static constexpr int kMaxStreams = 8;
  static constexpr size_t kArenaMaxBytes = size_t(16) << 20;
  DeviceTables() = default;
  DeviceTables(const DeviceTables &) = delete;
  DeviceTables &operator=(const DeviceTables &) = delete;
  ~DeviceTables() { (void)release(); }

  int upload(const void *host, size_t bytes);  // on the current device; `host` may go away after
  // The same in two steps, without a host image of the caller's: begin()
  // hands out `bytes` of host memory to fill (the pinned staging block), and
  // commit() makes it the tables' contents. data() is valid after commit().
  int begin(size_t bytes, char **host);
  int commit();
  int prepare(hipStream_t s);                  // before every launch that reads the tables
  int release();                               // idempotent
  const char *data() const { return static_cast<const char *>(ptr_); }
  int device() const { return device_; }

 private:
  void *ptr_ = nullptr;
  int device_ = -1;
  int cls_ = -1;              // arena size class of ptr_; -1: pool allocation
  size_t bytes_ = 0;
  void *staging_ = nullptr;   // pinned image until the copy is queued
  std::vector<char> image_;   // begin() above the arena's classes: filled, uploaded by commit()
  bool queued_ = false;       // copy queued (or done at upload)
  bool ready_ = false;        // copy known to be complete
  uint64_t ticket_ = 0;       // the arena's record of the queued copy
  std::mutex mu_;
  hipStream_t streams_[kMaxStreams] = {};
  int nstreams_ = 0;
  bool overflow_ = false;
};
} // namespace fs
