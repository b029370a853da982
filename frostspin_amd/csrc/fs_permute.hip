// Abelian permute executor: one gather kernel over the packed output blocks.
//
// Every output coefficient (new block b, row i, col j) looks up the old dense
// (row, col) it comes from as row_a[i] + col_b[j] (two int32 pairs built by
// the planner), then the old block / in-block position of that row and
// column. Writes are fully coalesced (consecutive threads write consecutive
// coefficients of one block); reads are gathers served mostly by L2.
// Semantics follow _numba_abelian_transpose (abelian_symmetric_tensor.py:
// 103-250) inverted into destination order; coefficients whose old block is
// missing read as zero exactly as the reference's zero-initialised blocks.
#include <cstring>
#include <memory>

#include "fs_hip.h"
#include "fs_plan.h"

namespace {

struct PermDev {
  const int32_t *old_row_block, *old_row_pos, *old_col_block, *old_col_pos;
  const int32_t *row_a, *col_b;
  const int64_t *old_off, *old_rs, *old_cs;
  const int64_t *new_off, *new_rowstart, *new_colstart, *new_shapes;
  const int64_t *tile_prefix;
  int32_t nblocks;
};

template <typename T>
__global__ void __launch_bounds__(256)
permute_gather_kernel(const T *__restrict__ src, T *__restrict__ dst, PermDev p) {
  const int64_t tile = blockIdx.x;
  int lo = 0, hi = p.nblocks;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (p.tile_prefix[mid] <= tile) lo = mid;
    else hi = mid;
  }
  const int b = lo;
  const int64_t nr = p.new_shapes[2 * b];
  const uint32_t nc = static_cast<uint32_t>(p.new_shapes[2 * b + 1]);
  const uint32_t n = static_cast<uint32_t>(nr * nc);
  const uint32_t base = static_cast<uint32_t>((tile - p.tile_prefix[b]) * kPermuteTile);
  const int2 *ra = reinterpret_cast<const int2 *>(p.row_a) + p.new_rowstart[b];
  const int2 *cb = reinterpret_cast<const int2 *>(p.col_b) + p.new_colstart[b];
  T *out = dst + p.new_off[b];
#pragma unroll 4
  for (int k = 0; k < static_cast<int>(kPermuteTile / 256); ++k) {
    const uint32_t e = base + static_cast<uint32_t>(k * 256) + threadIdx.x;
    if (e < n) {
      const uint32_t i = e / nc;
      const uint32_t j = e - i * nc;
      const int2 a = ra[i];
      const int2 c = cb[j];
      const int32_t R = a.x + c.x;
      const int32_t C = a.y + c.y;
      const int32_t ob = p.old_row_block[R];
      T v = T(0);
      if (ob >= 0 && ob == p.old_col_block[C]) {
        v = src[p.old_off[ob] + static_cast<int64_t>(p.old_row_pos[R]) * p.old_rs[ob] +
                static_cast<int64_t>(p.old_col_pos[C]) * p.old_cs[ob]];
      }
      out[e] = v;
    }
  }
}

template <typename V> size_t bytes_of(const std::vector<V> &v) { return v.size() * sizeof(V); }

int upload(fs_permute_plan *p) {
  int dev = 0;
  FS_HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(p->upload_mu);
  if (p->d_tables && p->device == dev) return FS_OK;
  if (p->d_tables) {
    fs::set_error("permute plan was uploaded on device %d, current device is %d", p->device, dev);
    return FS_ERR_INVALID;
  }
  // one allocation, 256-byte aligned sub-buffers
  struct Part { const void *src; size_t bytes; const void **dst; };
  std::vector<Part> parts = {
      {p->old_row_block.data(), bytes_of(p->old_row_block), (const void **)&p->d_old_row_block},
      {p->old_row_pos.data(), bytes_of(p->old_row_pos), (const void **)&p->d_old_row_pos},
      {p->old_col_block.data(), bytes_of(p->old_col_block), (const void **)&p->d_old_col_block},
      {p->old_col_pos.data(), bytes_of(p->old_col_pos), (const void **)&p->d_old_col_pos},
      {p->row_a.data(), bytes_of(p->row_a), (const void **)&p->d_row_a},
      {p->col_b.data(), bytes_of(p->col_b), (const void **)&p->d_col_b},
      {p->old_off.data(), bytes_of(p->old_off), (const void **)&p->d_old_off},
      {p->old_rs.data(), bytes_of(p->old_rs), (const void **)&p->d_old_rs},
      {p->old_cs.data(), bytes_of(p->old_cs), (const void **)&p->d_old_cs},
      {p->new_off.data(), bytes_of(p->new_off), (const void **)&p->d_new_off},
      {p->new_rowstart.data(), bytes_of(p->new_rowstart), (const void **)&p->d_new_rowstart},
      {p->new_colstart.data(), bytes_of(p->new_colstart), (const void **)&p->d_new_colstart},
      {p->new_shapes.data(), bytes_of(p->new_shapes), (const void **)&p->d_new_shapes},
      {p->tile_prefix.data(), bytes_of(p->tile_prefix), (const void **)&p->d_tile_prefix},
  };
  size_t total = 0;
  std::vector<size_t> offs;
  for (auto &q : parts) {
    offs.push_back(total);
    total += (q.bytes + 255) / 256 * 256 + 256;
  }
  std::vector<char> host(total, 0);
  for (size_t i = 0; i < parts.size(); ++i)
    if (parts[i].bytes) std::memcpy(host.data() + offs[i], parts[i].src, parts[i].bytes);
  auto tables = std::make_unique<fs::DeviceTables>();
  int st = tables->upload(host.data(), total);
  if (st != FS_OK) return st;
  for (size_t i = 0; i < parts.size(); ++i) *parts[i].dst = tables->data() + offs[i];
  p->d_tables = tables.release();
  p->device = dev;
  return FS_OK;
}

template <typename T>
int execute(fs_permute_plan *plan, const T *src, T *dst, void *stream) {
  FS_CHECK_ARG(plan, "permute execute: null plan");
  const int64_t nb = static_cast<int64_t>(plan->new_irreps.size());
  if (nb == 0 || plan->new_ncoeff == 0) return FS_OK;
  FS_CHECK_ARG(dst, "permute execute: null destination");
  FS_CHECK_ARG(src || plan->old_ncoeff == 0, "permute execute: null source");
  for (int64_t b = 0; b < nb; ++b)
    FS_CHECK_ARG(plan->new_shapes[2 * b] * plan->new_shapes[2 * b + 1] < (int64_t(1) << 32),
                 "permute execute: block too large for 32-bit in-block indexing");
  int st = upload(plan);
  if (st != FS_OK) return st;
  PermDev p;
  p.old_row_block = plan->d_old_row_block;
  p.old_row_pos = plan->d_old_row_pos;
  p.old_col_block = plan->d_old_col_block;
  p.old_col_pos = plan->d_old_col_pos;
  p.row_a = plan->d_row_a;
  p.col_b = plan->d_col_b;
  p.old_off = plan->d_old_off;
  p.old_rs = plan->d_old_rs;
  p.old_cs = plan->d_old_cs;
  p.new_off = plan->d_new_off;
  p.new_rowstart = plan->d_new_rowstart;
  p.new_colstart = plan->d_new_colstart;
  p.new_shapes = plan->d_new_shapes;
  p.tile_prefix = plan->d_tile_prefix;
  p.nblocks = static_cast<int32_t>(nb);
  const int64_t ntiles = plan->tile_prefix.back();
  FS_CHECK_ARG(ntiles < (int64_t(1) << 31), "permute execute: too many tiles");
  static_cast<fs::DeviceTables *>(plan->d_tables)->note_stream(fs::as_stream(stream));
  hipLaunchKernelGGL(permute_gather_kernel<T>, dim3(static_cast<unsigned>(ntiles)), dim3(256), 0,
                     fs::as_stream(stream), src, dst, p);
  FS_LAUNCH_CHECK("permute_gather_kernel");
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
