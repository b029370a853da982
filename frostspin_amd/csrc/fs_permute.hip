// Abelian permute executor: one workgroup per box tile (see fs_plan.h).
//
// A tile is a dense box of at most kPermuteTileVol coefficients whose source
// and destination positions are affine in the legs' digits. Load phase: the
// workgroup walks the tile in source order (fastest source digit first, so
// neighbouring lanes read neighbouring addresses along the longest contiguous
// source run) and drops each coefficient into LDS at its destination-order
// offset. Store phase: it walks the LDS copy in destination order and writes
// contiguous destination runs. Digit decoding uses one multiply-high per digit
// (q = umulhi(f, ceil(2^32 / len)) is exact for f < 2^12). Semantics follow
// _numba_abelian_transpose (abelian_symmetric_tensor.py:103-250); tiles of an
// absent old block write zeros, as the reference's zero-initialised blocks.
#include <cstring>
#include <memory>

#include "fs_hip.h"
#include "fs_plan.h"

namespace {

constexpr int kPermThreads = 512;

// Elements per thread handled together: their loads are issued back to back
// before any of them is stored, so each lane keeps several reads in flight.
constexpr int kPermBatch = 4;

template <typename T>
__global__ void __launch_bounds__(kPermThreads)
permute_box_kernel(const T *__restrict__ src, T *__restrict__ dst, const PermTile *__restrict__ tiles) {
  __shared__ T buf[kPermuteTileVol];
  const PermTile &t = tiles[blockIdx.x];
  const int vol = t.vol;
  T *out = dst + t.dst_base;
  if (t.src_base >= 0) {
    const T *in = src + t.src_base;
    const int ns = t.ns;
    for (int f0 = threadIdx.x; f0 < vol; f0 += kPermThreads * kPermBatch) {
      int32_t so[kPermBatch], lo[kPermBatch];
      T v[kPermBatch];
#pragma unroll
      for (int u = 0; u < kPermBatch; ++u) {
        const int f = f0 + u * kPermThreads;
        uint32_t rem = static_cast<uint32_t>(f < vol ? f : 0);
        int32_t a = 0, b = 0;
#pragma unroll
        for (int k = 0; k < kPermuteMaxDigits; ++k) {
          if (k < ns) {
            const uint32_t q = __umulhi(rem, t.smag[k]);
            const int32_t r = static_cast<int32_t>(rem - q * static_cast<uint32_t>(t.slen[k]));
            a += r * t.sstr[k];
            b += r * t.lstr[k];
            rem = q;
          }
        }
        so[u] = a;
        lo[u] = b;
      }
#pragma unroll
      for (int u = 0; u < kPermBatch; ++u)
        if (f0 + u * kPermThreads < vol) v[u] = in[so[u]];
#pragma unroll
      for (int u = 0; u < kPermBatch; ++u)
        if (f0 + u * kPermThreads < vol) buf[lo[u]] = v[u];
    }
  } else {
    for (int f = threadIdx.x; f < vol; f += kPermThreads) buf[f] = T(0);
  }
  __syncthreads();
  const int nd = t.nd;
  for (int f = threadIdx.x; f < vol; f += kPermThreads) {
    uint32_t rem = static_cast<uint32_t>(f);
    int32_t doff = 0;
#pragma unroll
    for (int k = 0; k < kPermuteMaxDigits; ++k) {
      if (k < nd) {
        const uint32_t q = __umulhi(rem, t.dmag[k]);
        const int32_t r = static_cast<int32_t>(rem - q * static_cast<uint32_t>(t.dlen[k]));
        doff += r * t.dstr[k];
        rem = q;
      }
    }
    out[doff] = buf[f];
  }
}

int upload(fs_permute_plan *p) {
  int dev = 0;
  FS_HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(p->upload_mu);
  if (p->d_tables && p->device == dev) return FS_OK;
  if (p->d_tables) {
    fs::set_error("permute plan was uploaded on device %d, current device is %d", p->device, dev);
    return FS_ERR_INVALID;
  }
  auto tables = std::make_unique<fs::DeviceTables>();
  int st = tables->upload(p->tiles.data(), p->tiles.size() * sizeof(PermTile));
  if (st != FS_OK) return st;
  p->d_tiles = reinterpret_cast<const PermTile *>(tables->data());
  p->d_tables = tables.release();
  p->device = dev;
  return FS_OK;
}

template <typename T>
int execute(fs_permute_plan *plan, const T *src, T *dst, void *stream) {
  FS_CHECK_ARG(plan, "permute execute: null plan");
  if (plan->tiles.empty() || plan->new_ncoeff == 0) return FS_OK;
  FS_CHECK_ARG(dst, "permute execute: null destination");
  FS_CHECK_ARG(src || plan->old_ncoeff == 0, "permute execute: null source");
  const int64_t ntiles = static_cast<int64_t>(plan->tiles.size());
  FS_CHECK_ARG(ntiles < (int64_t(1) << 31), "permute execute: too many tiles");
  int st = upload(plan);
  if (st != FS_OK) return st;
  static_cast<fs::DeviceTables *>(plan->d_tables)->note_stream(fs::as_stream(stream));
  hipLaunchKernelGGL(permute_box_kernel<T>, dim3(static_cast<unsigned>(ntiles)), dim3(kPermThreads), 0,
                     fs::as_stream(stream), src, dst, plan->d_tiles);
  FS_LAUNCH_CHECK("permute_box_kernel");
  return FS_OK;
}

} // namespace

extern "C" {

int fs_device_count(int *count) {
  return fs::guarded([&]() -> int {
    FS_CHECK_ARG(count, "fs_device_count: null pointer");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    (void)hipGetLastError();
    *count = n;
    return FS_OK;
  });
}

int fs_permute_plan_execute_f64(fs_permute_plan *plan, const double *src, double *dst,
                                void *stream) {
  return fs::guarded([&]() -> int { return execute<double>(plan, src, dst, stream); });
}

int fs_permute_plan_execute_f32(fs_permute_plan *plan, const float *src, float *dst,
                                void *stream) {
  return fs::guarded([&]() -> int { return execute<float>(plan, src, dst, stream); });
}

int fs_permute_plan_destroy(fs_permute_plan *plan) {
  return fs::guarded([&]() -> int {
    if (!plan) return FS_OK;
    auto *tables = static_cast<fs::DeviceTables *>(plan->d_tables);
    const int st = tables ? tables->release() : FS_OK;
    delete tables;
    delete plan;
    return st;
  });
}

} // extern "C"
