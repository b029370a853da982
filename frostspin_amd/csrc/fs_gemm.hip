// Grouped GEMM over uneven symmetry blocks on CDNA4 matrix cores.
//
// All blocks of one block-sparse contraction run in a single launch: the
// workgroup index is mapped to (group, tile_m, tile_n) through a prefix sum
// of per-group tile counts carried in the kernel arguments. Each workgroup
// (4 wavefronts of 64 lanes) computes a BM x BN tile with FP64 MFMA
// v_mfma_f64_16x16x4f64; operands are staged through LDS stored k-major so
// that the 16 lanes sharing one k read 16 consecutive doubles.
//
// Operand layouts are generic (row stride, column stride), so transposed
// (Fortran-ordered) blocks produced by dagger()/transpose() are consumed
// without a copy: the global->LDS load order follows whichever dimension is
// contiguous.
#include <algorithm>
#include <cstring>
#include <vector>

#include "fs_hip.h"

namespace {

constexpr int kMaxGroups = 32;
constexpr int kThreads = 256;

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

struct GroupDev {
  int32_t m, n, k, tiles_n;
  int32_t a_rs, a_cs, b_rs, b_cs;
  int32_t c_ld, pad;
  int64_t a_off, b_off, c_off;
};

struct Batch {
  const double *A;
  const double *B;
  double *C;
  double alpha, beta;
  int32_t ngroups, xcd_remap;
  int32_t tile_prefix[kMaxGroups + 1];
  GroupDev g[kMaxGroups];
};

// Map a linear tile id to its group (linear scan over <= 32 prefix entries).
__device__ inline int find_group(const Batch &bt, int tile) {
  int g = 0;
  while (g + 1 < bt.ngroups && bt.tile_prefix[g + 1] <= tile) ++g;
  return g;
}

// Launch-order remap so that consecutive logical tiles land on the same XCD
// (workgroups are dealt round-robin to the 8 XCDs): tiles that share an A row
// panel then share one L2.
__device__ inline int xcd_logical_tile(int bid, int nblocks) {
  const int nx = 8;
  const int per = nblocks / nx;
  const int full = per * nx;
  if (bid >= full) return bid;
  return (bid % nx) * per + bid / nx;
}

template <int BM, int BN, int BK>
__global__ void __launch_bounds__(kThreads) gemm_f64_kernel(Batch bt) {
  // LDS tiles, k-major: As[k][i], Bs[k][j]; +1 pad breaks power-of-two strides.
  __shared__ double As[BK][BM + 1];
  __shared__ double Bs[BK][BN + 1];

  const int tile = bt.xcd_remap ? xcd_logical_tile(blockIdx.x, gridDim.x) : blockIdx.x;
  const int gi = find_group(bt, tile);
  const GroupDev &G = bt.g[gi];
  const int t = tile - bt.tile_prefix[gi];
  const int tm = t / G.tiles_n;
  const int tn = t - tm * G.tiles_n;
  const int m0 = tm * BM;
  const int n0 = tn * BN;
  const double *__restrict__ A = bt.A + G.a_off;
  const double *__restrict__ B = bt.B + G.b_off;

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  // 4 waves in a 2x2 arrangement; each wave owns (BM/2) x (BN/2)
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int TI = WM / 16, TJ = WN / 16;
  const int wm = wave >> 1, wn = wave & 1;

  d4 acc[TI][TJ];
#pragma unroll
  for (int a = 0; a < TI; ++a)
#pragma unroll
    for (int b = 0; b < TJ; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};

  const bool a_kcontig = (G.a_cs == 1);
  const bool b_ncontig = (G.b_cs == 1);
  constexpr int A_PER = BM * BK / kThreads;
  constexpr int B_PER = BN * BK / kThreads;

  for (int k0 = 0; k0 < G.k; k0 += BK) {
    double ra[A_PER], rb[B_PER];
#pragma unroll
    for (int r = 0; r < A_PER; ++r) {
      const int idx = threadIdx.x + kThreads * r;
      const int i = a_kcontig ? idx / BK : idx % BM;
      const int kk = a_kcontig ? idx % BK : idx / BM;
      const int gi_ = m0 + i, gk = k0 + kk;
      ra[r] = (gi_ < G.m && gk < G.k)
                  ? A[static_cast<int64_t>(gi_) * G.a_rs + static_cast<int64_t>(gk) * G.a_cs]
                  : 0.0;
    }
#pragma unroll
    for (int r = 0; r < B_PER; ++r) {
      const int idx = threadIdx.x + kThreads * r;
      const int j = b_ncontig ? idx % BN : idx / BK;
      const int kk = b_ncontig ? idx / BN : idx % BK;
      const int gj = n0 + j, gk = k0 + kk;
      rb[r] = (gj < G.n && gk < G.k)
                  ? B[static_cast<int64_t>(gk) * G.b_rs + static_cast<int64_t>(gj) * G.b_cs]
                  : 0.0;
    }
    __syncthreads();  // previous tile fully consumed
#pragma unroll
    for (int r = 0; r < A_PER; ++r) {
      const int idx = threadIdx.x + kThreads * r;
      const int i = a_kcontig ? idx / BK : idx % BM;
      const int kk = a_kcontig ? idx % BK : idx / BM;
      As[kk][i] = ra[r];
    }
#pragma unroll
    for (int r = 0; r < B_PER; ++r) {
      const int idx = threadIdx.x + kThreads * r;
      const int j = b_ncontig ? idx % BN : idx / BK;
      const int kk = b_ncontig ? idx / BN : idx % BK;
      Bs[kk][j] = rb[r];
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      const int kr = ks + (lane >> 4);
      double av[TI], bv[TJ];
#pragma unroll
      for (int a = 0; a < TI; ++a) av[a] = As[kr][wm * WM + a * 16 + (lane & 15)];
#pragma unroll
      for (int b = 0; b < TJ; ++b) bv[b] = Bs[kr][wn * WN + b * 16 + (lane & 15)];
#pragma unroll
      for (int a = 0; a < TI; ++a)
#pragma unroll
        for (int b = 0; b < TJ; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
  }

  // epilogue: for v_mfma_f64_16x16x4f64 lane l holds D[l/16 + 4*r][l%16] of
  // each 16x16 tile (rows interleaved by 4, measured on gfx950)
  double *__restrict__ C = bt.C + G.c_off;
  const double alpha = bt.alpha, beta = bt.beta;
#pragma unroll
  for (int a = 0; a < TI; ++a)
#pragma unroll
    for (int b = 0; b < TJ; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * WM + a * 16 + (lane >> 4) + 4 * r;
        const int col = n0 + wn * WN + b * 16 + (lane & 15);
        if (row < G.m && col < G.n) {
          double *c = C + static_cast<int64_t>(row) * G.c_ld + col;
          const double v = alpha * acc[a][b][r];
          *c = (beta == 0.0) ? v : v + beta * (*c);
        }
      }
}

// ---------------------------------------------------------------------------
// Compensated FP32 form: x = hi + lo with hi = fp32(x), lo = fp32(x - hi);
// A*B ~ hi_a*hi_b + hi_a*lo_b + lo_a*hi_b on v_mfma_f32_16x16x4f32 with FP32
// accumulation inside one BK chunk and FP64 accumulation across chunks.
// ---------------------------------------------------------------------------
template <int BM, int BN, int BK>
__global__ void __launch_bounds__(kThreads) gemm_split_kernel(Batch bt) {
  __shared__ float Ah[BK][BM + 1], Al[BK][BM + 1];
  __shared__ float Bh[BK][BN + 1], Bl[BK][BN + 1];

  const int tile = bt.xcd_remap ? xcd_logical_tile(blockIdx.x, gridDim.x) : blockIdx.x;
  const int gi = find_group(bt, tile);
  const GroupDev &G = bt.g[gi];
  const int t = tile - bt.tile_prefix[gi];
  const int tm = t / G.tiles_n;
  const int tn = t - tm * G.tiles_n;
  const int m0 = tm * BM;
  const int n0 = tn * BN;
  const double *__restrict__ A = bt.A + G.a_off;
  const double *__restrict__ B = bt.B + G.b_off;

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int TI = WM / 16, TJ = WN / 16;
  const int wm = wave >> 1, wn = wave & 1;

  d4 acc[TI][TJ];
#pragma unroll
  for (int a = 0; a < TI; ++a)
#pragma unroll
    for (int b = 0; b < TJ; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};

  const bool a_kcontig = (G.a_cs == 1);
  const bool b_ncontig = (G.b_cs == 1);
  constexpr int A_PER = BM * BK / kThreads;
  constexpr int B_PER = BN * BK / kThreads;

  for (int k0 = 0; k0 < G.k; k0 += BK) {
    double ra[A_PER], rb[B_PER];
#pragma unroll
    for (int r = 0; r < A_PER; ++r) {
      const int idx = threadIdx.x + kThreads * r;
      const int i = a_kcontig ? idx / BK : idx % BM;
      const int kk = a_kcontig ? idx % BK : idx / BM;
      const int gi_ = m0 + i, gk = k0 + kk;
      ra[r] = (gi_ < G.m && gk < G.k)
                  ? A[static_cast<int64_t>(gi_) * G.a_rs + static_cast<int64_t>(gk) * G.a_cs]
                  : 0.0;
    }
#pragma unroll
    for (int r = 0; r < B_PER; ++r) {
      const int idx = threadIdx.x + kThreads * r;
      const int j = b_ncontig ? idx % BN : idx / BK;
      const int kk = b_ncontig ? idx / BN : idx % BK;
      const int gj = n0 + j, gk = k0 + kk;
      rb[r] = (gj < G.n && gk < G.k)
                  ? B[static_cast<int64_t>(gk) * G.b_rs + static_cast<int64_t>(gj) * G.b_cs]
                  : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < A_PER; ++r) {
      const int idx = threadIdx.x + kThreads * r;
      const int i = a_kcontig ? idx / BK : idx % BM;
      const int kk = a_kcontig ? idx % BK : idx / BM;
      const float h = static_cast<float>(ra[r]);
      Ah[kk][i] = h;
      Al[kk][i] = static_cast<float>(ra[r] - static_cast<double>(h));
    }
#pragma unroll
    for (int r = 0; r < B_PER; ++r) {
      const int idx = threadIdx.x + kThreads * r;
      const int j = b_ncontig ? idx % BN : idx / BK;
      const int kk = b_ncontig ? idx / BN : idx % BK;
      const float h = static_cast<float>(rb[r]);
      Bh[kk][j] = h;
      Bl[kk][j] = static_cast<float>(rb[r] - static_cast<double>(h));
    }
    __syncthreads();
    f4 part[TI][TJ];
#pragma unroll
    for (int a = 0; a < TI; ++a)
#pragma unroll
      for (int b = 0; b < TJ; ++b) part[a][b] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      const int kr = ks + (lane >> 4);
      float ah[TI], al[TI], bh[TJ], bl[TJ];
#pragma unroll
      for (int a = 0; a < TI; ++a) {
        ah[a] = Ah[kr][wm * WM + a * 16 + (lane & 15)];
        al[a] = Al[kr][wm * WM + a * 16 + (lane & 15)];
      }
#pragma unroll
      for (int b = 0; b < TJ; ++b) {
        bh[b] = Bh[kr][wn * WN + b * 16 + (lane & 15)];
        bl[b] = Bl[kr][wn * WN + b * 16 + (lane & 15)];
      }
#pragma unroll
      for (int a = 0; a < TI; ++a)
#pragma unroll
        for (int b = 0; b < TJ; ++b) {
          part[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(al[a], bh[b], part[a][b], 0, 0, 0);
          part[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah[a], bl[b], part[a][b], 0, 0, 0);
          part[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah[a], bh[b], part[a][b], 0, 0, 0);
        }
    }
#pragma unroll
    for (int a = 0; a < TI; ++a)
#pragma unroll
      for (int b = 0; b < TJ; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[a][b][r] += static_cast<double>(part[a][b][r]);
  }

  double *__restrict__ C = bt.C + G.c_off;
  const double alpha = bt.alpha, beta = bt.beta;
#pragma unroll
  for (int a = 0; a < TI; ++a)
#pragma unroll
    for (int b = 0; b < TJ; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * WM + a * 16 + 4 * (lane >> 4) + r;
        const int col = n0 + wn * WN + b * 16 + (lane & 15);
        if (row < G.m && col < G.n) {
          double *c = C + static_cast<int64_t>(row) * G.c_ld + col;
          const double v = alpha * acc[a][b][r];
          *c = (beta == 0.0) ? v : v + beta * (*c);
        }
      }
}

constexpr int kBM = 64, kBN = 64, kBK = 16;

template <bool SPLIT>
int launch_grouped(int ngroups, const fs_gemm_group *groups, const double *A, const double *B,
                   double *C, double alpha, double beta, void *stream) {
  FS_CHECK_ARG(ngroups >= 0, "grouped gemm: negative group count");
  FS_CHECK_ARG(ngroups == 0 || groups, "grouped gemm: null groups");
  int g0 = 0;
  while (g0 < ngroups) {
    Batch bt;
    std::memset(&bt, 0, sizeof(bt));
    bt.A = A;
    bt.B = B;
    bt.C = C;
    bt.alpha = alpha;
    bt.beta = beta;
    int ng = 0;
    int64_t tiles = 0;
    bt.tile_prefix[0] = 0;
    for (; g0 < ngroups && ng < kMaxGroups; ++g0) {
      const fs_gemm_group &q = groups[g0];
      FS_CHECK_ARG(q.m >= 0 && q.n >= 0 && q.k >= 0, "grouped gemm: negative dimension");
      if (q.m == 0 || q.n == 0) continue;
      FS_CHECK_ARG(q.m < (1ll << 31) && q.n < (1ll << 31) && q.k < (1ll << 31),
                   "grouped gemm: dimension too large");
      FS_CHECK_ARG(A && B && C, "grouped gemm: null operand");
      GroupDev &d = bt.g[ng];
      d.m = static_cast<int32_t>(q.m);
      d.n = static_cast<int32_t>(q.n);
      d.k = static_cast<int32_t>(q.k);
      d.tiles_n = static_cast<int32_t>(fs::ceil_div(q.n, kBN));
      d.a_rs = static_cast<int32_t>(q.a_rs);
      d.a_cs = static_cast<int32_t>(q.a_cs);
      d.b_rs = static_cast<int32_t>(q.b_rs);
      d.b_cs = static_cast<int32_t>(q.b_cs);
      d.c_ld = static_cast<int32_t>(q.c_ld);
      d.a_off = q.a_off;
      d.b_off = q.b_off;
      d.c_off = q.c_off;
      tiles += fs::ceil_div(q.m, kBM) * d.tiles_n;
      FS_CHECK_ARG(tiles < (1ll << 31), "grouped gemm: too many tiles");
      bt.tile_prefix[ng + 1] = static_cast<int32_t>(tiles);
      ++ng;
    }
    if (ng == 0) continue;
    bt.ngroups = ng;
    bt.xcd_remap = tiles >= 64 ? 1 : 0;
    if (SPLIT)
      hipLaunchKernelGGL((gemm_split_kernel<kBM, kBN, kBK>), dim3(static_cast<unsigned>(tiles)),
                         dim3(kThreads), 0, fs::as_stream(stream), bt);
    else
      hipLaunchKernelGGL((gemm_f64_kernel<kBM, kBN, kBK>), dim3(static_cast<unsigned>(tiles)),
                         dim3(kThreads), 0, fs::as_stream(stream), bt);
    FS_LAUNCH_CHECK(SPLIT ? "gemm_split_kernel" : "gemm_f64_kernel");
  }
  return FS_OK;
}

} // namespace

extern "C" {

int fs_grouped_gemm_f64(int ngroups, const fs_gemm_group *groups, const double *A,
                        const double *B, double *C, double alpha, double beta, void *stream) {
  return fs::guarded(
      [&]() -> int { return launch_grouped<false>(ngroups, groups, A, B, C, alpha, beta, stream); });
}

int fs_grouped_gemm_f64_split(int ngroups, const fs_gemm_group *groups, const double *A,
                              const double *B, double *C, double alpha, double beta,
                              void *stream) {
  return fs::guarded(
      [&]() -> int { return launch_grouped<true>(ngroups, groups, A, B, C, alpha, beta, stream); });
}

} // extern "C"
