// Plan-table memory: uploads and releases that never synchronise a compute
// stream (fs::DeviceTables, declared in fs_hip.h).
#include <cstdint>
#include <vector>

#include "fs_hip.h"

namespace fs {
namespace {

std::mutex g_upload_mu;
std::vector<hipStream_t> g_upload_streams;  // per device, created on first use

// The device's upload stream. On creation the default pool of the device is
// told to keep freed memory (release threshold unlimited), so trimming the pool
// never returns memory to the driver at a synchronisation point, and not to
// reuse a block by making the allocating stream wait for the stream that freed
// it: a plan freed on a compute stream would otherwise make the next upload
// wait for that stream's queued kernels. Blocks whose free has completed are
// still reused; otherwise the pool grows.
int upload_stream(int dev, hipStream_t *out) {
  std::lock_guard<std::mutex> lk(g_upload_mu);
  if (static_cast<int>(g_upload_streams.size()) <= dev) g_upload_streams.resize(dev + 1, nullptr);
  if (!g_upload_streams[dev]) {
    hipStream_t s = nullptr;
    FS_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipMemPool_t pool = nullptr;
    if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
      uint64_t keep = UINT64_MAX;
      FS_HIP_CHECK(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
      int no = 0;
      FS_HIP_CHECK(hipMemPoolSetAttribute(pool, hipMemPoolReuseAllowInternalDependencies, &no));
    } else {
      (void)hipGetLastError();
    }
    g_upload_streams[dev] = s;
  }
  *out = g_upload_streams[dev];
  return FS_OK;
}

} // namespace

int DeviceTables::upload(const void *host, size_t bytes) {
  FS_CHECK_ARG(ptr_ == nullptr, "plan tables uploaded twice");
  int dev = 0;
  FS_HIP_CHECK(hipGetDevice(&dev));
  hipStream_t us = nullptr;
  int st = upload_stream(dev, &us);
  if (st != FS_OK) return st;
  void *p = nullptr;
  FS_HIP_CHECK(hipMallocAsync(&p, bytes ? bytes : 256, us));
  hipError_t e = bytes ? hipMemcpyAsync(p, host, bytes, hipMemcpyHostToDevice, us) : hipSuccess;
  if (e == hipSuccess) e = hipStreamSynchronize(us);  // the copy only: `host` may go away
  if (e != hipSuccess) {
    (void)hipFreeAsync(p, us);
    set_error("upload of plan tables failed: %s", hipGetErrorString(e));
    return FS_ERR_HIP;
  }
  ptr_ = p;
  device_ = dev;
  return FS_OK;
}

void DeviceTables::note_stream(hipStream_t s) {
  std::lock_guard<std::mutex> lk(mu_);
  for (int i = 0; i < nstreams_; ++i)
    if (streams_[i] == s) return;
  if (nstreams_ < kMaxStreams) streams_[nstreams_++] = s;
  else overflow_ = true;
}

int DeviceTables::release() {
  std::lock_guard<std::mutex> lk(mu_);
  if (!ptr_) return FS_OK;
  int cur = 0;
  FS_HIP_CHECK(hipGetDevice(&cur));
  if (cur != device_) FS_HIP_CHECK(hipSetDevice(device_));
  hipError_t e = hipSuccess;
  if (overflow_) {
    e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipFreeAsync(ptr_, nullptr);
  } else if (nstreams_ == 1) {
    e = hipFreeAsync(ptr_, streams_[0]);  // after the last launch queued there
  } else {
    hipStream_t us = nullptr;
    int st = upload_stream(device_, &us);
    if (st != FS_OK) {
      if (cur != device_) (void)hipSetDevice(cur);
      return st;
    }
    // the upload stream waits for the queued work of every stream that read
    // the tables, then frees them
    for (int i = 0; i < nstreams_ && e == hipSuccess; ++i) {
      hipEvent_t ev = nullptr;
      e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
      if (e == hipSuccess) e = hipEventRecord(ev, streams_[i]);
      if (e == hipSuccess) e = hipStreamWaitEvent(us, ev, 0);
      if (ev) (void)hipEventDestroy(ev);
    }
    if (e == hipSuccess) e = hipFreeAsync(ptr_, us);
  }
  if (cur != device_) (void)hipSetDevice(cur);
  ptr_ = nullptr;
  nstreams_ = 0;
  overflow_ = false;
  if (e != hipSuccess) {
    set_error("release of plan tables failed: %s", hipGetErrorString(e));
    return FS_ERR_HIP;
  }
  return FS_OK;
}

} // namespace fs
