// Plan-table memory: uploads and releases that never synchronise a compute
// stream (fs::DeviceTables, declared in fs_hip.h).
#include <cstdint>
#include <cstring>
#include <deque>
#include <memory>
#include <vector>

#include "fs_hip.h"

namespace fs {
namespace {

std::mutex g_upload_mu;
std::vector<hipStream_t> g_upload_streams;  // per device, created on first use

// The device's upload stream (tables above the arena's largest class). On
// creation the default pool of the device is told to keep freed memory (release
// threshold unlimited), so trimming the pool never returns memory to the driver
// at a synchronisation point, and not to reuse a block by making the allocating
// stream wait for the stream that freed it: a plan freed on a compute stream
// would otherwise make the next upload wait for that stream's queued kernels.
// Blocks whose free has completed are still reused; otherwise the pool grows.
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

// ---------------------------------------------------------------------------
// Per-device arena of table blocks. Class c holds blocks of kMinBlock << c
// bytes; device and pinned host blocks are carved from slabs of at least
// kSlabBytes that are never returned. A block goes back to its free list only
// after the events recorded behind its last use have completed: a staging
// block after its host-to-device copy, a device block after the work queued on
// every stream that executed its plan. Both waiting lists are polled front to
// back (oldest first) on every upload and release; polling stops at the first
// record still in flight.

constexpr size_t kMinBlock = 256;
constexpr size_t kSlabBytes = size_t(2) << 20;
constexpr int kClasses = 17;  // 256 B .. 16 MiB
static_assert((kMinBlock << (kClasses - 1)) == DeviceTables::kArenaMaxBytes, "arena classes");

struct InFlight {
  hipEvent_t ev[DeviceTables::kMaxStreams] = {};
  int nev = 0;
  void *block = nullptr;
  int cls = 0;
  uint64_t ticket = 0;  // copies only: consecutive from 1
};

struct Arena {
  std::mutex mu;
  std::vector<void *> free_dev[kClasses], free_host[kClasses];
  std::deque<InFlight> copies;  // staging blocks behind queued copies, in ticket order
  std::deque<InFlight> frees;   // device blocks behind released tables
  uint64_t next_ticket = 1;
  uint64_t done_ticket = 0;     // every copy up to this ticket has completed
  std::vector<hipEvent_t> events;
};

std::vector<std::unique_ptr<Arena>> g_arenas;  // per device, guarded by g_upload_mu

Arena &arena_of(int dev) {
  std::lock_guard<std::mutex> lk(g_upload_mu);
  if (static_cast<int>(g_arenas.size()) <= dev) g_arenas.resize(dev + 1);
  if (!g_arenas[dev]) g_arenas[dev] = std::make_unique<Arena>();
  return *g_arenas[dev];
}

int class_of(size_t bytes) {
  int c = 0;
  while ((kMinBlock << c) < bytes) ++c;
  return c < kClasses ? c : -1;
}

// 1: complete, 0: in flight, <0: error (reported)
int query(hipEvent_t ev) {
  const hipError_t e = hipEventQuery(ev);
  if (e == hipSuccess) return 1;
  (void)hipGetLastError();  // a query that is not ready must not read as a launch failure
  if (e == hipErrorNotReady) return 0;
  set_error("plan tables: event query failed: %s", hipGetErrorString(e));
  return -1;
}

bool all_done(const InFlight &f) {
  for (int i = 0; i < f.nev; ++i)
    if (query(f.ev[i]) != 1) return false;
  return true;
}

// with the arena locked, on its device
void poll(Arena &a) {
  while (!a.copies.empty() && all_done(a.copies.front())) {
    InFlight &f = a.copies.front();
    a.free_host[f.cls].push_back(f.block);
    for (int i = 0; i < f.nev; ++i) a.events.push_back(f.ev[i]);
    a.done_ticket = f.ticket;
    a.copies.pop_front();
  }
  while (!a.frees.empty() && all_done(a.frees.front())) {
    InFlight &f = a.frees.front();
    a.free_dev[f.cls].push_back(f.block);
    for (int i = 0; i < f.nev; ++i) a.events.push_back(f.ev[i]);
    a.frees.pop_front();
  }
}

int take_event(Arena &a, hipEvent_t *ev) {
  if (!a.events.empty()) {
    *ev = a.events.back();
    a.events.pop_back();
    return FS_OK;
  }
  FS_HIP_CHECK(hipEventCreateWithFlags(ev, hipEventDisableTiming));
  return FS_OK;
}

int take_block(Arena &a, int cls, bool host, void **out) {
  std::vector<void *> &list = host ? a.free_host[cls] : a.free_dev[cls];
  if (list.empty()) {
    const size_t block = kMinBlock << cls;
    const size_t slab = block > kSlabBytes ? block : kSlabBytes;
    void *p = nullptr;
    if (host) FS_HIP_CHECK(hipHostMalloc(&p, slab, hipHostMallocDefault));
    else FS_HIP_CHECK(hipMalloc(&p, slab));
    for (size_t off = slab; off >= block; off -= block) list.push_back(static_cast<char *>(p) + off - block);
  }
  *out = list.back();
  list.pop_back();
  return FS_OK;
}

} // namespace

int DeviceTables::begin(size_t bytes, char **host) {
  std::lock_guard<std::mutex> lk(mu_);
  FS_CHECK_ARG(ptr_ == nullptr && image_.empty(), "plan tables uploaded twice");
  *host = nullptr;
  int dev = 0;
  FS_HIP_CHECK(hipGetDevice(&dev));
  const int cls = class_of(bytes);
  if (cls < 0) {
    // above the arena: filled in host memory, copied through the pool by commit()
    image_.resize(bytes);
    *host = image_.data();
    device_ = dev;
    bytes_ = bytes;
    return FS_OK;
  }
  Arena &a = arena_of(dev);
  void *d = nullptr, *h = nullptr;
  {
    std::lock_guard<std::mutex> al(a.mu);
    poll(a);
    int st = take_block(a, cls, false, &d);
    if (st == FS_OK && bytes) st = take_block(a, cls, true, &h);
    if (st != FS_OK) {
      if (d) a.free_dev[cls].push_back(d);
      return st;
    }
  }
  ptr_ = d;
  staging_ = h;
  cls_ = cls;
  bytes_ = bytes;
  device_ = dev;
  queued_ = ready_ = bytes == 0;
  *host = static_cast<char *>(h);
  return FS_OK;
}

int DeviceTables::commit() {
  std::lock_guard<std::mutex> lk(mu_);
  if (image_.empty()) return FS_OK;  // arena block: the copy is queued by the first prepare()
  hipStream_t us = nullptr;
  int st = upload_stream(device_, &us);
  if (st != FS_OK) return st;
  void *p = nullptr;
  FS_HIP_CHECK(hipMallocAsync(&p, bytes_, us));
  hipError_t e = hipMemcpyAsync(p, image_.data(), bytes_, hipMemcpyHostToDevice, us);
  if (e == hipSuccess) e = hipStreamSynchronize(us);  // the copy only: the image goes away
  if (e != hipSuccess) {
    (void)hipFreeAsync(p, us);
    set_error("upload of plan tables failed: %s", hipGetErrorString(e));
    return FS_ERR_HIP;
  }
  std::vector<char>().swap(image_);
  ptr_ = p;
  cls_ = -1;
  queued_ = ready_ = true;
  return FS_OK;
}

int DeviceTables::upload(const void *host, size_t bytes) {
  char *h = nullptr;
  int st = begin(bytes, &h);
  if (st != FS_OK) return st;
  if (bytes) std::memcpy(h, host, bytes);
  return commit();
}

int DeviceTables::prepare(hipStream_t s) {
  std::lock_guard<std::mutex> lk(mu_);
  FS_CHECK_ARG(ptr_ != nullptr, "plan tables used after release");
  for (int i = 0; i < nstreams_; ++i)
    if (streams_[i] == s) return FS_OK;  // already ordered after the copy
  // A graph recorded on `s` must not hold the tables' copy, or a wait for it,
  // with the arena's events behind them: plans are built and first used
  // outside a capture (a later use on another stream needs neither once the
  // copy is known to be complete).
  auto refuse_capture = [&]() -> int {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    FS_HIP_CHECK(hipStreamIsCapturing(s, &cs));
    FS_CHECK_ARG(cs == hipStreamCaptureStatusNone,
                 "plan tables: first use of a plan on a stream that is being captured");
    return FS_OK;
  };
  if (!queued_) {
    int st = refuse_capture();
    if (st != FS_OK) return st;
    // first use: the copy goes on this stream, ahead of the plan's kernels
    Arena &a = arena_of(device_);
    std::lock_guard<std::mutex> al(a.mu);
    InFlight f;
    st = take_event(a, &f.ev[0]);
    if (st != FS_OK) return st;
    hipError_t e = hipMemcpyAsync(ptr_, staging_, bytes_, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
      e = hipEventRecord(f.ev[0], s);
      if (e != hipSuccess) {
        // the copy may be in flight: keep the staging block out of reuse
        (void)hipStreamSynchronize(s);
        staging_ = nullptr;
      }
    }
    if (e != hipSuccess) {
      a.events.push_back(f.ev[0]);
      set_error("plan tables: copy to the device failed: %s", hipGetErrorString(e));
      return FS_ERR_HIP;
    }
    f.nev = 1;
    f.block = staging_;
    f.cls = cls_;
    f.ticket = ticket_ = a.next_ticket++;
    a.copies.push_back(f);
    staging_ = nullptr;
    queued_ = true;
  } else if (!ready_) {
    // the copy was queued on another stream: wait for it unless it is done
    Arena &a = arena_of(device_);
    std::lock_guard<std::mutex> al(a.mu);
    if (ticket_ <= a.done_ticket) {
      ready_ = true;
    } else {
      // not known to be complete: querying the copy's event (recorded on
      // another stream, which may be capturing now) is refused under capture
      int st = refuse_capture();
      if (st != FS_OK) return st;
      const InFlight &f = a.copies[static_cast<size_t>(ticket_ - a.copies.front().ticket)];
      const int q = query(f.ev[0]);
      if (q < 0) return FS_ERR_HIP;
      if (q == 1) ready_ = true;
      else FS_HIP_CHECK(hipStreamWaitEvent(s, f.ev[0], 0));
    }
  }
  if (nstreams_ < kMaxStreams) streams_[nstreams_++] = s;
  else overflow_ = true;
  return FS_OK;
}

int DeviceTables::release() {
  std::lock_guard<std::mutex> lk(mu_);
  if (!ptr_) return FS_OK;
  int cur = 0;
  FS_HIP_CHECK(hipGetDevice(&cur));
  if (cur != device_) FS_HIP_CHECK(hipSetDevice(device_));
  hipError_t e = hipSuccess;
  int st = FS_OK;
  if (cls_ >= 0) {
    Arena &a = arena_of(device_);
    if (queued_ && overflow_) e = hipDeviceSynchronize();
    std::lock_guard<std::mutex> al(a.mu);
    if (staging_) a.free_host[cls_].push_back(staging_);  // never copied
    if (!queued_ || overflow_ || nstreams_ == 0) {
      if (e == hipSuccess) a.free_dev[cls_].push_back(ptr_);  // nothing queued reads it
    } else {
      // reused once the work queued so far on every stream that read it is done
      InFlight f;
      f.block = ptr_;
      f.cls = cls_;
      for (int i = 0; i < nstreams_ && st == FS_OK && e == hipSuccess; ++i) {
        st = take_event(a, &f.ev[i]);
        if (st != FS_OK) break;
        e = hipEventRecord(f.ev[i], streams_[i]);
        if (e != hipSuccess) a.events.push_back(f.ev[i]);
        else f.nev = i + 1;
      }
      if (st == FS_OK && e == hipSuccess) {
        a.frees.push_back(f);
      } else {
        for (int i = 0; i < f.nev; ++i) a.events.push_back(f.ev[i]);
        // the block is dropped rather than reused while kernels may read it
      }
      poll(a);
    }
  } else if (overflow_) {
    e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipFreeAsync(ptr_, nullptr);
  } else if (nstreams_ <= 1) {
    e = hipFreeAsync(ptr_, nstreams_ ? streams_[0] : nullptr);  // after the last launch queued there
  } else {
    hipStream_t us = nullptr;
    st = upload_stream(device_, &us);
    // the upload stream waits for the queued work of every stream that read
    // the tables, then frees them
    for (int i = 0; st == FS_OK && i < nstreams_ && e == hipSuccess; ++i) {
      hipEvent_t ev = nullptr;
      e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
      if (e == hipSuccess) e = hipEventRecord(ev, streams_[i]);
      if (e == hipSuccess) e = hipStreamWaitEvent(us, ev, 0);
      if (ev) (void)hipEventDestroy(ev);
    }
    if (st == FS_OK && e == hipSuccess) e = hipFreeAsync(ptr_, us);
  }
  if (cur != device_) (void)hipSetDevice(cur);
  ptr_ = nullptr;
  staging_ = nullptr;
  nstreams_ = 0;
  overflow_ = false;
  if (st != FS_OK) return st;
  if (e != hipSuccess) {
    set_error("release of plan tables failed: %s", hipGetErrorString(e));
    return FS_ERR_HIP;
  }
  return FS_OK;
}

} // namespace fs

extern "C" int fs_plan_tables_settle(void) {
  return fs::guarded([&]() -> int {
    std::vector<fs::Arena *> arenas;
    {
      std::lock_guard<std::mutex> lk(fs::g_upload_mu);
      for (auto &a : fs::g_arenas)
        if (a) arenas.push_back(a.get());
    }
    for (fs::Arena *a : arenas) {
      std::lock_guard<std::mutex> al(a->mu);
      fs::poll(*a);
    }
    return FS_OK;
  });
}
