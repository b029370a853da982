// Abelian permute executor: one workgroup per tile of a box (see fs_plan.h).
//
// A tile is a dense box slice of at most kPermuteTileVol coefficients whose
// source and destination positions are affine in the legs' digits. The
// workgroup finds its box (one table lookup), derives its tile from the box
// record (tile_of_box: slow-digit indices and chunk of the cut digit, the last
// chunk's shorter length), then loads: it walks the tile in source order
// (fastest source digit first, so neighbouring lanes read neighbouring
// addresses along the longest contiguous source run) and drops each
// coefficient into LDS at its destination-order offset. Store phase: it walks
// the LDS copy in destination order and writes contiguous destination runs.
// Digit decoding uses one multiply-high per digit (q = umulhi(f, ceil(2^32 /
// len)) is exact for f < 2^12). Semantics follow _numba_abelian_transpose
// (abelian_symmetric_tensor.py:103-250); tiles of an absent old block write
// zeros, as the reference's zero-initialised blocks.
#include <cstring>
#include <memory>

#include "fs_hip.h"
#include "fs_plan.h"

namespace {

constexpr int kPermThreads = 512;
constexpr int K = kPermuteMaxDigits;

// Elements per thread handled together: their loads are issued back to back
// before any of them is stored, so each lane keeps several reads in flight
// (8: a full 4096-element tile in flight per 512-thread workgroup).
constexpr int kPermBatch = 8;

// One tile of a box, as the kernel walks it (and as fs_permute_plan_tiles
// reports it).
struct TileView {
  int64_t src_base, dst_base;  // src_base < 0: zeros
  int32_t vol, ns, nd;
  uint32_t smag[K];
  int32_t slen[K], sstr[K], lstr[K];
  uint32_t dmag[K];
  int32_t dlen[K], dstr[K];
};

// Tile `local` of box b: chunk of the cut digit fastest, then the slow digits
// (fastest first); the last chunk patches (or drops) one source entry. Every
// array index is a constant after unrolling, so the tile stays in registers.
__host__ __device__ inline void tile_of_box(const PermBox &b, int32_t local, TileView &t) {
  const int32_t c = local % b.nchunks;
  int32_t r = local / b.nchunks;
  int64_t so = static_cast<int64_t>(c) * b.chunk * b.cut_s;
  int64_t d = b.dst_base + static_cast<int64_t>(c) * b.chunk * b.cut_d;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (k < b.nslow) {
      const int32_t idx = r % b.slow_len[k];
      r /= b.slow_len[k];
      so += idx * b.slow_s[k];
      d += idx * b.slow_d[k];
    }
  }
  t.src_base = b.src_base >= 0 ? b.src_base + so : -1;
  t.dst_base = d;
  const bool edge = b.nchunks > 1 && c == b.nchunks - 1;
  t.vol = edge ? b.vol_edge : b.vol;
  const bool patch = edge && b.s_cut >= 0;
  const bool drop = patch && b.s_drop_edge;
  t.ns = drop ? b.ns - 1 : b.ns;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    constexpr int kLast = K - 1;
    const int n1 = k < kLast ? k + 1 : kLast;  // dropped entry: the slower ones move down
    const bool shift = drop && k >= b.s_cut;
    t.slen[k] = shift ? b.slen[n1] : b.slen[k];
    t.smag[k] = shift ? b.smag[n1] : b.smag[k];
    t.sstr[k] = shift ? b.sstr[n1] : b.sstr[k];
    t.lstr[k] = shift ? b.lstr[n1] : b.lstr[k];
    if (patch && !drop && k == b.s_cut) {
      t.slen[k] = b.s_len_edge;
      t.smag[k] = b.s_mag_edge;
    }
  }
  t.nd = b.nd;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    t.dlen[k] = b.dlen[k];
    t.dmag[k] = b.dmag[k];
    t.dstr[k] = b.dstr[k];
    // the last destination entry holds the cut digit: rem < its length
    // there, so its multiply-high gives 0 with either magic
    if (edge && k == b.nd - 1) t.dlen[k] = b.d_len_edge;
  }
}

// Workgroups are dealt round-robin to the 8 XCDs, each with its own L2.
// Consecutive tiles (neighbouring boxes of one source block, sharing the
// cache lines at their edges) go to the same XCD: workgroup b runs tile
// (b mod 8) * (n / 8) + b / 8 (the last n mod 8 tiles in order).
__device__ inline int32_t xcd_tile(int32_t b, int32_t n) {
  constexpr int32_t kXcds = 8;
  const int32_t per = n / kXcds;
  return b >= per * kXcds ? b : (b % kXcds) * per + b / kXcds;
}

template <typename T>
__global__ void __launch_bounds__(kPermThreads)
permute_box_kernel(const T *__restrict__ src, T *__restrict__ dst, const PermBox *__restrict__ boxes,
                   const int32_t *__restrict__ tile0, const int32_t *__restrict__ tile_box) {
  __shared__ T buf[kPermuteTileVol];
  const int32_t tile = xcd_tile(static_cast<int32_t>(blockIdx.x), static_cast<int32_t>(gridDim.x));
  const int32_t bi = tile_box[tile];
  TileView t;
  tile_of_box(boxes[bi], tile - tile0[bi], t);
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
  // one buffer: boxes, then tile0, then tile_box (each 256-byte aligned)
  const size_t nb = p->boxes.size();
  const size_t off_t0 = fs::align_up(nb * sizeof(PermBox), 256);
  const size_t off_tb = off_t0 + fs::align_up(p->tile0.size() * sizeof(int32_t), 256);
  const size_t bytes = off_tb + p->tile_box.size() * sizeof(int32_t);
  char *host = nullptr;  // staging memory, filled in place
  int st = tables->begin(bytes, &host);
  if (st != FS_OK) return st;
  std::memset(host, 0, bytes);
  if (nb) std::memcpy(host, p->boxes.data(), nb * sizeof(PermBox));
  std::memcpy(host + off_t0, p->tile0.data(), p->tile0.size() * sizeof(int32_t));
  std::memcpy(host + off_tb, p->tile_box.data(), p->tile_box.size() * sizeof(int32_t));
  st = tables->commit();
  if (st != FS_OK) return st;
  p->d_boxes = reinterpret_cast<const PermBox *>(tables->data());
  p->d_tile0 = reinterpret_cast<const int32_t *>(tables->data() + off_t0);
  p->d_tile_box = reinterpret_cast<const int32_t *>(tables->data() + off_tb);
  p->d_tables = tables.release();
  p->device = dev;
  return FS_OK;
}

template <typename T>
int execute(fs_permute_plan *plan, const T *src, T *dst, void *stream) {
  FS_CHECK_ARG(plan, "permute execute: null plan");
  if (plan->ntiles == 0 || plan->new_ncoeff == 0) return FS_OK;
  FS_CHECK_ARG(dst, "permute execute: null destination");
  FS_CHECK_ARG(src || plan->old_ncoeff == 0, "permute execute: null source");
  const int64_t ntiles = plan->ntiles;
  FS_CHECK_ARG(ntiles < (int64_t(1) << 31), "permute execute: too many tiles");
  int st = upload(plan);
  if (st == FS_OK) st = static_cast<fs::DeviceTables *>(plan->d_tables)->prepare(fs::as_stream(stream));
  if (st != FS_OK) return st;
  hipLaunchKernelGGL(permute_box_kernel<T>, dim3(static_cast<unsigned>(ntiles)), dim3(kPermThreads), 0,
                     fs::as_stream(stream), src, dst, plan->d_boxes, plan->d_tile0,
                     plan->d_tile_box);
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

int fs_permute_plan_tiles(const fs_permute_plan *plan, int64_t *ntiles, int64_t *nboxes,
                          int64_t *out) {
  return fs::guarded([&]() -> int {
    FS_CHECK_ARG(plan && ntiles && nboxes, "fs_permute_plan_tiles: null argument");
    *ntiles = plan->ntiles;
    *nboxes = static_cast<int64_t>(plan->boxes.size());
    if (!out) return FS_OK;
    static_assert(FS_PERMUTE_TILE_RECORD == 5 + 5 * K, "tile record layout");
    TileView t;
    for (int64_t i = 0; i < plan->ntiles; ++i) {  // the kernel's own derivation of each tile
      const int32_t bi = plan->tile_box[i];
      tile_of_box(plan->boxes[bi], static_cast<int32_t>(i - plan->tile0[bi]), t);
      int64_t *r = out;
      std::memset(r, 0, FS_PERMUTE_TILE_RECORD * sizeof(int64_t));
      r[0] = t.src_base;
      r[1] = t.dst_base;
      r[2] = t.vol;
      r[3] = t.ns;
      r[4] = t.nd;
      for (int k = 0; k < t.ns; ++k) {
        r[5 + 3 * k] = t.slen[k];
        r[6 + 3 * k] = t.sstr[k];
        r[7 + 3 * k] = t.lstr[k];
      }
      for (int k = 0; k < t.nd; ++k) {
        r[5 + 3 * K + 2 * k] = t.dlen[k];
        r[6 + 3 * K + 2 * k] = t.dstr[k];
      }
      out += FS_PERMUTE_TILE_RECORD;
    }
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
