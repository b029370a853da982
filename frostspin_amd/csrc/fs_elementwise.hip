// Bandwidth-bound helpers on packed block buffers: norms, scaling, axpby,
// batched strided 2-D copies (transposes, slicing, diagonal products),
// traces, deterministic random fill and precision conversion.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "fs_hip.h"

namespace {

constexpr int kThreads = 256;
constexpr int kMaxSeg = 64;
constexpr int kNormChunk = 4096;  // elements per chunk (16 per thread)
constexpr int kNormGrid = 1024;   // partial sums (fixed => deterministic)

struct NormArgs {
  const double *x;
  double *partial;
  int32_t nseg, pad;
  int64_t seg_off[kMaxSeg + 1];
  int64_t chunk_prefix[kMaxSeg + 1];
  double w[kMaxSeg];
};

__device__ inline double block_sum(double v, double *red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kThreads / 64; ++i) s += red[i];
  __syncthreads();
  return s;  // valid in thread 0
}

__global__ void __launch_bounds__(kThreads) norm2_partial_kernel(NormArgs a) {
  __shared__ double red[kThreads / 64];
  const int64_t nchunks = a.chunk_prefix[a.nseg];
  double acc = 0.0;
  for (int64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
    int s = 0;
    while (s + 1 < a.nseg && a.chunk_prefix[s + 1] <= c) ++s;
    const int64_t start = a.seg_off[s] + (c - a.chunk_prefix[s]) * kNormChunk;
    const int64_t end = min(start + static_cast<int64_t>(kNormChunk), a.seg_off[s + 1]);
    double loc = 0.0;
    for (int64_t e = start + threadIdx.x; e < end; e += kThreads) {
      const double v = a.x[e];
      loc += v * v;
    }
    acc += a.w[s] * loc;
  }
  const double tot = block_sum(acc, red);
  if (threadIdx.x == 0) a.partial[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kThreads) sum_partials_kernel(const double *partial, int n,
                                                                double *out, int accumulate) {
  __shared__ double red[kThreads / 64];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += kThreads) acc += partial[i];
  const double tot = block_sum(acc, red);
  if (threadIdx.x == 0) out[0] = accumulate ? out[0] + tot : tot;
}

__global__ void scale_kernel(int64_t n, double alpha, double *x) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    x[i] *= alpha;
}

__global__ void scale_rsqrt_dev_kernel(int64_t n, const double *d, double *x) {
  const double f = 1.0 / sqrt(d[0]);
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    x[i] *= f;
}

__global__ void axpby_kernel(int64_t n, double alpha, const double *x, double beta, double *y) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    y[i] = (beta == 0.0) ? alpha * x[i] : alpha * x[i] + beta * y[i];
}

constexpr int kMaxCopy = 32;
constexpr int kCopyTile = 2048;

struct CopyDev {
  int64_t rows, cols;
  int64_t s_off, s_rs, s_cs;
  int64_t d_off, d_rs, d_cs;
  int64_t rw_off, cw_off, ci_off;
  double alpha, beta;
};

struct CopyBatch {
  const double *src;
  double *dst;
  const double *row_w;
  const double *col_w;
  const int64_t *col_idx;
  int32_t n, pad;
  int64_t tile_prefix[kMaxCopy + 1];
  CopyDev d[kMaxCopy];
};

// Flat element order follows the destination when it is row-contiguous,
// otherwise the source, so at least one side is coalesced.
__global__ void __launch_bounds__(kThreads) copy2d_kernel(CopyBatch b) {
  const int64_t tile = blockIdx.x;
  int g = 0;
  while (g + 1 < b.n && b.tile_prefix[g + 1] <= tile) ++g;
  const CopyDev D = b.d[g];
  const int64_t base = (tile - b.tile_prefix[g]) * kCopyTile;
  const int64_t n = D.rows * D.cols;
  const bool col_major_walk = (D.d_rs == 1 && D.d_cs != 1) || (D.d_cs != 1 && D.s_rs == 1);
  for (int k = 0; k < kCopyTile / kThreads; ++k) {
    const int64_t e = base + k * kThreads + threadIdx.x;
    if (e >= n) break;
    int64_t i, j;
    if (col_major_walk) {
      j = e / D.rows;
      i = e - j * D.rows;
    } else {
      i = e / D.cols;
      j = e - i * D.cols;
    }
    const int64_t jj = b.col_idx ? b.col_idx[D.ci_off + j] : j;
    double v = D.alpha * b.src[D.s_off + i * D.s_rs + jj * D.s_cs];
    if (b.row_w) v *= b.row_w[D.rw_off + i];
    if (b.col_w) v *= b.col_w[D.cw_off + j];
    double *d = b.dst + D.d_off + i * D.d_rs + j * D.d_cs;
    *d = (D.beta == 0.0) ? v : v + D.beta * (*d);
  }
}

struct TraceArgs {
  const double *x;
  double *out;
  int32_t nseg, pad;
  int64_t off[kMaxSeg], n[kMaxSeg], ld[kMaxSeg];
  double w[kMaxSeg];
};

__global__ void __launch_bounds__(kThreads) trace_kernel(TraceArgs a) {
  __shared__ double red[kThreads / 64];
  double acc = 0.0;
  for (int s = 0; s < a.nseg; ++s) {
    double loc = 0.0;
    for (int64_t i = threadIdx.x; i < a.n[s]; i += kThreads) loc += a.x[a.off[s] + i * (a.ld[s] + 1)];
    acc += a.w[s] * loc;
  }
  const double tot = block_sum(acc, red);
  if (threadIdx.x == 0) a.out[0] = tot;
}

__device__ inline uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Box-Muller on two 53-bit uniforms derived from a hash of (seed, pair index).
__global__ void fill_normal_kernel(int64_t n, uint64_t seed, double *x) {
  const double two_pi = 6.283185307179586476925286766559;
  for (int64_t pidx = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; 2 * pidx < n;
       pidx += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const uint64_t h1 = splitmix64(seed ^ splitmix64(2 * static_cast<uint64_t>(pidx)));
    const uint64_t h2 = splitmix64(seed ^ splitmix64(2 * static_cast<uint64_t>(pidx) + 1));
    const double u1 = (static_cast<double>(h1 >> 11) + 1.0) * (1.0 / 9007199254740992.0);
    const double u2 = static_cast<double>(h2 >> 11) * (1.0 / 9007199254740992.0);
    const double r = sqrt(-2.0 * log(u1));
    x[2 * pidx] = r * cos(two_pi * u2);
    if (2 * pidx + 1 < n) x[2 * pidx + 1] = r * sin(two_pi * u2);
  }
}

__global__ void f64_to_f32_kernel(int64_t n, const double *x, float *y) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    y[i] = static_cast<float>(x[i]);
}

__global__ void f32_to_f64_kernel(int64_t n, const float *x, double *y) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    y[i] = static_cast<double>(x[i]);
}

inline unsigned grid_for(int64_t n) {
  const int64_t g = fs::ceil_div(n, kThreads * 4);
  return static_cast<unsigned>(g < 1 ? 1 : (g > 8192 ? 8192 : g));
}

} // namespace

namespace {

// Two-term weighted gather, the data movement of O(2) tensors: every output
// element combines at most 2 source rows and 2 source columns,
//   out[i][j] = beta*out[i][j] + sum_{a,b<2} rw[i][a] cw[j][b] src[ri[i][a]][ci[j][b]],
// which covers the Sz = 0 block split into 0odd / 0even (fixed points with
// weight 1, doublets with +-1/sqrt(2); o2_symmetric_tensor.py:118-310), its
// inverse, and the reflected Sz < 0 blocks (one term, sign weight;
// :325-352, :798-851). Terms with zero weight are skipped.
constexpr int kMaxGather = 32;

struct GatherDev {
  int64_t rows, cols;
  int64_t s_off, s_rs, s_cs;
  int64_t d_off, d_rs, d_cs;
  int64_t ri_off, rw_off, ci_off, cw_off;
  double beta;
};

struct GatherBatch {
  const double *src;
  double *dst;
  const int64_t *idx;
  const double *w;
  int32_t n, pad;
  int64_t tile_prefix[kMaxGather + 1];
  GatherDev d[kMaxGather];
};

__global__ void __launch_bounds__(kThreads) gather2d_kernel(GatherBatch b) {
  const int64_t tile = blockIdx.x;
  int g = 0;
  while (g + 1 < b.n && b.tile_prefix[g + 1] <= tile) ++g;
  const GatherDev D = b.d[g];
  const int64_t base = (tile - b.tile_prefix[g]) * kCopyTile;
  const int64_t n = D.rows * D.cols;
  for (int k = 0; k < kCopyTile / kThreads; ++k) {
    const int64_t e = base + k * kThreads + threadIdx.x;
    if (e >= n) break;
    const int64_t i = e / D.cols;
    const int64_t j = e - i * D.cols;
    double v = 0.0;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const double wr = b.w[D.rw_off + 2 * i + a];
      if (wr == 0.0) continue;
      const int64_t r = b.idx[D.ri_off + 2 * i + a];
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const double wc = b.w[D.cw_off + 2 * j + c];
        if (wc == 0.0) continue;
        v = fma(wr * wc, b.src[D.s_off + r * D.s_rs + b.idx[D.ci_off + 2 * j + c] * D.s_cs], v);
      }
    }
    double *d = b.dst + D.d_off + i * D.d_rs + j * D.d_cs;
    *d = (D.beta == 0.0) ? v : v + D.beta * (*d);
  }
}

int check_gather(int ndesc, const fs_gather2d_desc *descs, const int64_t *idx, int64_t nidx,
                 int64_t nw, int64_t src_size, int64_t dst_size) {
  FS_CHECK_ARG(ndesc >= 0 && nidx >= 0 && nw >= 0, "gather2d: negative size");
  FS_CHECK_ARG(ndesc == 0 || descs, "gather2d: null descriptors");
  for (int t = 0; t < ndesc; ++t) {
    const fs_gather2d_desc &q = descs[t];
    FS_CHECK_ARG(q.rows >= 0 && q.cols >= 0 && q.src_rows >= 0 && q.src_cols >= 0,
                 "gather2d: negative extent");
    if (q.rows == 0 || q.cols == 0) continue;
    FS_CHECK_ARG(idx != nullptr, "gather2d: null index table");
    FS_CHECK_ARG(q.src_rows > 0 && q.src_cols > 0, "gather2d: empty source for a non-empty output");
    FS_CHECK_ARG(q.s_off >= 0 && q.s_rs >= 0 && q.s_cs >= 0 && q.d_off >= 0 && q.d_rs >= 0 && q.d_cs >= 0,
                 "gather2d: negative offset or stride");
    FS_CHECK_ARG(q.s_off + (q.src_rows - 1) * q.s_rs + (q.src_cols - 1) * q.s_cs < src_size,
                 "gather2d: source out of bounds");
    FS_CHECK_ARG(q.d_off + (q.rows - 1) * q.d_rs + (q.cols - 1) * q.d_cs < dst_size,
                 "gather2d: destination out of bounds");
    FS_CHECK_ARG(q.ri_off >= 0 && q.ri_off + 2 * q.rows <= nidx && q.ci_off >= 0 &&
                     q.ci_off + 2 * q.cols <= nidx,
                 "gather2d: index table out of range");
    FS_CHECK_ARG(q.rw_off >= 0 && q.rw_off + 2 * q.rows <= nw && q.cw_off >= 0 &&
                     q.cw_off + 2 * q.cols <= nw,
                 "gather2d: weight table out of range");
    for (int64_t i = 0; i < 2 * q.rows; ++i)
      FS_CHECK_ARG(idx[q.ri_off + i] >= 0 && idx[q.ri_off + i] < q.src_rows, "gather2d: row index out of range");
    for (int64_t j = 0; j < 2 * q.cols; ++j)
      FS_CHECK_ARG(idx[q.ci_off + j] >= 0 && idx[q.ci_off + j] < q.src_cols, "gather2d: column index out of range");
  }
  return FS_OK;
}

} // namespace

extern "C" int fs_gather2d_check(int ndesc, const fs_gather2d_desc *descs, const int64_t *idx,
                                 int64_t nidx, int64_t nw, int64_t src_size, int64_t dst_size) {
  return fs::guarded([&]() -> int { return check_gather(ndesc, descs, idx, nidx, nw, src_size, dst_size); });
}

extern "C" int fs_batched_gather2d_f64(int ndesc, const fs_gather2d_desc *descs, const int64_t *idx,
                                       const double *w, const double *src, double *dst,
                                       void *stream) {
  return fs::guarded([&]() -> int {
    FS_CHECK_ARG(ndesc >= 0 && (ndesc == 0 || descs), "gather2d: bad descriptors");
    int i0 = 0;
    while (i0 < ndesc) {
      GatherBatch b;
      std::memset(&b, 0, sizeof(b));
      b.src = src;
      b.dst = dst;
      b.idx = idx;
      b.w = w;
      int nb = 0;
      int64_t tiles = 0;
      for (; i0 < ndesc && nb < kMaxGather; ++i0) {
        const fs_gather2d_desc &q = descs[i0];
        FS_CHECK_ARG(q.rows >= 0 && q.cols >= 0, "gather2d: negative extent");
        if (q.rows == 0 || q.cols == 0) continue;
        FS_CHECK_ARG(src && dst && idx && w, "gather2d: null buffer");
        GatherDev &d = b.d[nb];
        d.rows = q.rows;
        d.cols = q.cols;
        d.s_off = q.s_off;
        d.s_rs = q.s_rs;
        d.s_cs = q.s_cs;
        d.d_off = q.d_off;
        d.d_rs = q.d_rs;
        d.d_cs = q.d_cs;
        d.ri_off = q.ri_off;
        d.rw_off = q.rw_off;
        d.ci_off = q.ci_off;
        d.cw_off = q.cw_off;
        d.beta = q.beta;
        b.tile_prefix[nb] = tiles;
        tiles += fs::ceil_div(q.rows * q.cols, kCopyTile);
        ++nb;
      }
      if (nb == 0) continue;
      b.tile_prefix[nb] = tiles;
      b.n = nb;
      FS_CHECK_ARG(tiles < (1ll << 31), "gather2d: too many tiles");
      hipLaunchKernelGGL(gather2d_kernel, dim3(static_cast<unsigned>(tiles)), dim3(kThreads), 0,
                         fs::as_stream(stream), b);
      FS_LAUNCH_CHECK("gather2d_kernel");
    }
    return FS_OK;
  });
}

extern "C" {

int fs_weighted_norm2_f64(int nseg, const int64_t *seg_off, const double *w, const double *x,
                          double *partial, double *out, void *stream) {
  return fs::guarded([&]() -> int {
    FS_CHECK_ARG(nseg >= 0 && out && partial, "weighted norm2: invalid arguments");
    FS_CHECK_ARG(nseg == 0 || (seg_off && w), "weighted norm2: null segment data");
    int first = 1;
    int s0 = 0;
    do {
      NormArgs a;
      std::memset(&a, 0, sizeof(a));
      a.x = x;
      a.partial = partial;
      const int ns = std::min(kMaxSeg, nseg - s0);
      a.nseg = ns;
      a.chunk_prefix[0] = 0;
      for (int s = 0; s < ns; ++s) {
        a.seg_off[s] = seg_off[s0 + s];
        a.seg_off[s + 1] = seg_off[s0 + s + 1];
        FS_CHECK_ARG(a.seg_off[s + 1] >= a.seg_off[s], "weighted norm2: decreasing offsets");
        a.w[s] = w[s0 + s];
        a.chunk_prefix[s + 1] =
            a.chunk_prefix[s] + fs::ceil_div(a.seg_off[s + 1] - a.seg_off[s], kNormChunk);
      }
      FS_CHECK_ARG(ns == 0 || a.chunk_prefix[ns] == 0 || x, "weighted norm2: null data");
      hipLaunchKernelGGL(norm2_partial_kernel, dim3(kNormGrid), dim3(kThreads), 0,
                         fs::as_stream(stream), a);
      FS_LAUNCH_CHECK("norm2_partial_kernel");
      hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(kThreads), 0, fs::as_stream(stream),
                         static_cast<const double *>(partial), kNormGrid, out, first ? 0 : 1);
      FS_LAUNCH_CHECK("sum_partials_kernel");
      first = 0;
      s0 += ns;
    } while (s0 < nseg);
    return FS_OK;
  });
}

int fs_scale_f64(int64_t n, double alpha, double *x, void *stream) {
  return fs::guarded([&]() -> int {
    if (n <= 0) return FS_OK;
    FS_CHECK_ARG(x, "scale: null pointer");
    hipLaunchKernelGGL(scale_kernel, dim3(grid_for(n)), dim3(kThreads), 0, fs::as_stream(stream),
                       n, alpha, x);
    FS_LAUNCH_CHECK("scale_kernel");
    return FS_OK;
  });
}

int fs_scale_rsqrt_dev_f64(int64_t n, const double *d, double *x, void *stream) {
  return fs::guarded([&]() -> int {
    if (n <= 0) return FS_OK;
    FS_CHECK_ARG(x && d, "scale_rsqrt_dev: null pointer");
    hipLaunchKernelGGL(scale_rsqrt_dev_kernel, dim3(grid_for(n)), dim3(kThreads), 0,
                       fs::as_stream(stream), n, d, x);
    FS_LAUNCH_CHECK("scale_rsqrt_dev_kernel");
    return FS_OK;
  });
}

int fs_axpby_f64(int64_t n, double alpha, const double *x, double beta, double *y, void *stream) {
  return fs::guarded([&]() -> int {
    if (n <= 0) return FS_OK;
    FS_CHECK_ARG(x && y, "axpby: null pointer");
    hipLaunchKernelGGL(axpby_kernel, dim3(grid_for(n)), dim3(kThreads), 0, fs::as_stream(stream),
                       n, alpha, x, beta, y);
    FS_LAUNCH_CHECK("axpby_kernel");
    return FS_OK;
  });
}

int fs_batched_copy2d_f64(int ndesc, const fs_copy2d_desc *descs, const double *src, double *dst,
                          const double *row_w, const double *col_w, const int64_t *col_idx,
                          void *stream) {
  return fs::guarded([&]() -> int {
    FS_CHECK_ARG(ndesc >= 0, "copy2d: negative count");
    FS_CHECK_ARG(ndesc == 0 || descs, "copy2d: null descriptors");
    int i0 = 0;
    while (i0 < ndesc) {
      CopyBatch b;
      std::memset(&b, 0, sizeof(b));
      b.src = src;
      b.dst = dst;
      b.row_w = row_w;
      b.col_w = col_w;
      b.col_idx = col_idx;
      int nb = 0;
      int64_t tiles = 0;
      for (; i0 < ndesc && nb < kMaxCopy; ++i0) {
        const fs_copy2d_desc &q = descs[i0];
        FS_CHECK_ARG(q.rows >= 0 && q.cols >= 0, "copy2d: negative extent");
        if (q.rows == 0 || q.cols == 0) continue;
        FS_CHECK_ARG(src && dst, "copy2d: null buffer");
        CopyDev &d = b.d[nb];
        d.rows = q.rows;
        d.cols = q.cols;
        d.s_off = q.s_off;
        d.s_rs = q.s_rs;
        d.s_cs = q.s_cs;
        d.d_off = q.d_off;
        d.d_rs = q.d_rs;
        d.d_cs = q.d_cs;
        d.rw_off = q.rw_off;
        d.cw_off = q.cw_off;
        d.ci_off = q.ci_off;
        d.alpha = q.alpha;
        d.beta = q.beta;
        b.tile_prefix[nb] = tiles;
        tiles += fs::ceil_div(q.rows * q.cols, kCopyTile);
        ++nb;
      }
      if (nb == 0) continue;
      b.tile_prefix[nb] = tiles;
      b.n = nb;
      FS_CHECK_ARG(tiles < (1ll << 31), "copy2d: too many tiles");
      hipLaunchKernelGGL(copy2d_kernel, dim3(static_cast<unsigned>(tiles)), dim3(kThreads), 0,
                         fs::as_stream(stream), b);
      FS_LAUNCH_CHECK("copy2d_kernel");
    }
    return FS_OK;
  });
}

int fs_weighted_trace_f64(int nseg, const int64_t *off, const int64_t *n, const int64_t *ld,
                          const double *w, const double *x, double *out, void *stream) {
  return fs::guarded([&]() -> int {
    FS_CHECK_ARG(nseg >= 0 && nseg <= kMaxSeg && out, "trace: invalid arguments (nseg <= 64)");
    TraceArgs a;
    std::memset(&a, 0, sizeof(a));
    a.x = x;
    a.out = out;
    a.nseg = nseg;
    for (int s = 0; s < nseg; ++s) {
      a.off[s] = off[s];
      a.n[s] = n[s];
      a.ld[s] = ld[s];
      a.w[s] = w[s];
    }
    hipLaunchKernelGGL(trace_kernel, dim3(1), dim3(kThreads), 0, fs::as_stream(stream), a);
    FS_LAUNCH_CHECK("trace_kernel");
    return FS_OK;
  });
}

int fs_fill_normal_f64(int64_t n, uint64_t seed, double *x, void *stream) {
  return fs::guarded([&]() -> int {
    if (n <= 0) return FS_OK;
    FS_CHECK_ARG(x, "fill_normal: null pointer");
    hipLaunchKernelGGL(fill_normal_kernel, dim3(grid_for((n + 1) / 2)), dim3(kThreads), 0,
                       fs::as_stream(stream), n, static_cast<unsigned long long>(seed), x);
    FS_LAUNCH_CHECK("fill_normal_kernel");
    return FS_OK;
  });
}

int fs_convert_f64_to_f32(int64_t n, const double *x, float *y, void *stream) {
  return fs::guarded([&]() -> int {
    if (n <= 0) return FS_OK;
    FS_CHECK_ARG(x && y, "convert: null pointer");
    hipLaunchKernelGGL(f64_to_f32_kernel, dim3(grid_for(n)), dim3(kThreads), 0,
                       fs::as_stream(stream), n, x, y);
    FS_LAUNCH_CHECK("f64_to_f32_kernel");
    return FS_OK;
  });
}

int fs_memcpy_async(void *dst, const void *src, int64_t bytes, void *stream) {
  return fs::guarded([&]() -> int {
    if (bytes <= 0) return FS_OK;
    FS_CHECK_ARG(dst && src, "memcpy: null pointer");
    FS_HIP_CHECK(hipMemcpyAsync(dst, src, static_cast<size_t>(bytes), hipMemcpyDefault,
                                fs::as_stream(stream)));
    return FS_OK;
  });
}

int fs_convert_f32_to_f64(int64_t n, const float *x, double *y, void *stream) {
  return fs::guarded([&]() -> int {
    if (n <= 0) return FS_OK;
    FS_CHECK_ARG(x && y, "convert: null pointer");
    hipLaunchKernelGGL(f32_to_f64_kernel, dim3(grid_for(n)), dim3(kThreads), 0,
                       fs::as_stream(stream), n, x, y);
    FS_LAUNCH_CHECK("f32_to_f64_kernel");
    return FS_OK;
  });
}

} // extern "C"
