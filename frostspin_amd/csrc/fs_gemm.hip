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
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "fs_hip.h"

namespace {

constexpr int kMaxGroups = 32;
constexpr int kThreads = 256;

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// Each group carries its own operand base addresses, so one launch can take
// its blocks from several tensors (the multi-operand entry point).
struct GroupDev {
  int32_t m, n, k, tiles_n;
  int32_t a_rs, a_cs, b_rs, b_cs;
  int32_t c_ld, kc;    // kc: K range of one split (multiple of the K panel)
  int32_t splits, pad;  // split-K: workgroups per output tile of this group
  const double *a;
  const double *b;
  double *c;
  int64_t ws_off;  // split-K: partial s of this group's m x n tile at ws + ws_off + s * m * n
};

struct Batch {
  double alpha, beta;
  int32_t ngroups, xcd_remap;
  int32_t tile_prefix[kMaxGroups + 1];
  GroupDev g[kMaxGroups];
  double *ws;  // null: no split-K, the epilogue writes C
  int64_t elem_prefix[kMaxGroups + 1];
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

// Global -> register tile loads: the index walk follows whichever dimension of
// the operand is contiguous so every wavefront load is coalesced. Loads are
// unconditional, from indices clamped into the matrix (m, n, k >= 1): a load
// under a branch, or a select right behind it, makes the compiler wait for
// every load in flight, which would serialise the prefetched K panels. Rows
// and columns past the matrix then hold copies of its last row or column and
// only reach output elements that are never written; K indices past the
// split's range are zeroed when the panel is stored to LDS (k_mask_a/b).
template <int BM, int BK, int PER, int NT = kThreads>
__device__ inline void load_a(double (&r)[PER], const double *__restrict__ A, const GroupDev &G,
                              int m0, int k0, bool kcontig) {
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int idx = threadIdx.x + NT * q;
    const int i = kcontig ? idx / BK : idx % BM;
    const int kk = kcontig ? idx % BK : idx / BM;
    r[q] = A[static_cast<int64_t>(min(m0 + i, G.m - 1)) * G.a_rs +
             static_cast<int64_t>(min(k0 + kk, G.k - 1)) * G.a_cs];
  }
}

template <int BN, int BK, int PER, int NT = kThreads>
__device__ inline void load_b(double (&r)[PER], const double *__restrict__ B, const GroupDev &G,
                              int n0, int k0, bool ncontig) {
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int idx = threadIdx.x + NT * q;
    const int j = ncontig ? idx % BN : idx / BK;
    const int kk = ncontig ? idx / BN : idx % BK;
    r[q] = B[static_cast<int64_t>(min(k0 + kk, G.k - 1)) * G.b_rs +
             static_cast<int64_t>(min(n0 + j, G.n - 1)) * G.b_cs];
  }
}

// FP64 MFMA tile kernel. BM x BN output tile per workgroup (4 wavefronts in a
// 2x2 arrangement, each (BM/2) x (BN/2) = TI x TJ MFMA 16x16 tiles), BK = 16.
// Software pipeline: the next K panel is fetched into registers while the
// current one is consumed from LDS; LDS is double buffered so each K step needs
// one workgroup barrier.
//
// LDS layout follows the source layout so that both the stores and the MFMA
// operand reads are bank-conflict free (64 banks x 4 B = 32 doubles):
//  * dimension-contiguous source (col-major A / row-major B): k-major rows
//    [k][i] with row stride BM+16 (= 16 mod 32): the 4 k rows read by one MFMA
//    operand land in disjoint bank halves, stores are unit stride;
//  * k-contiguous source (row-major A / col-major B): i-major rows [i][k] with
//    odd stride BK+1: the 16 rows of an operand read map to distinct banks and
//    the 4 k columns shift them by one bank.
// Occupancy target per tile shape: 64 x 64 tiles fit 4 workgroups per CU in
// LDS and, held to 128 registers (accumulators in VGPRs), 4 waves per SIMD
// (3 otherwise): 35.7 vs 32.5 TFLOP/s on the projector products of a SU(2)
// chi=256 move, 51 vs 45 on 4096^3. 32-deep K panels (2 workgroups per CU)
// measured slower. The other shapes keep the compiler's choice.
// A 128 x 128 tile with 8 waves (2 x 4, 16 instead of 8 flop per byte of
// L2 traffic) measured slower on every shape (45 vs 51 TFLOP/s at 4096^3).
template <int BM, int BN, int BK, int WGM, int WGN>
constexpr int gemm_waves_per_eu() {
  return (BM == 64 && BN == 64 && BK == 16 && WGM * WGN == 4) ? 4 : 1;
}

// PF: K panels in flight per workgroup (register prefetch depth). A launch
// with few workgroups per CU and a long K (thin products) is bound by the
// memory latency of its panel loads; PF panels in flight multiply its
// bandwidth. The panel loads of a K step are issued right after the previous
// panel went from registers to LDS, PF steps ahead of their use.
template <int BM, int BN, int BK, int WGM = 2, int WGN = 2, int PF = 1>
__global__ void __launch_bounds__(64 * WGM * WGN)
__attribute__((amdgpu_waves_per_eu(gemm_waves_per_eu<BM, BN, BK, WGM, WGN>()))) gemm_f64_kernel(Batch bt) {
  constexpr int NT = 64 * WGM * WGN;
  constexpr int PAD = 16;
  constexpr int A_SZ = (BK * (BM + PAD) > BM * (BK + 1)) ? BK * (BM + PAD) : BM * (BK + 1);
  constexpr int B_SZ = (BK * (BN + PAD) > BN * (BK + 1)) ? BK * (BN + PAD) : BN * (BK + 1);
  __shared__ double As[2][A_SZ];
  __shared__ double Bs[2][B_SZ];

  const int tile = bt.xcd_remap ? xcd_logical_tile(blockIdx.x, gridDim.x) : blockIdx.x;
  const int gi = find_group(bt, tile);
  const GroupDev &G = bt.g[gi];
  // split-K: a group with fewer splits than the launch leaves the rest idle
  // (whole workgroup, before any barrier)
  if (static_cast<int>(blockIdx.y) >= G.splits) return;
  const int t = tile - bt.tile_prefix[gi];
  const int tm = t / G.tiles_n;
  const int tn = t - tm * G.tiles_n;
  const int m0 = tm * BM;
  const int n0 = tn * BN;
  const double *__restrict__ A = G.a;
  const double *__restrict__ B = G.b;

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  constexpr int WM = BM / WGM, WN = BN / WGN;
  constexpr int TI = WM / 16, TJ = WN / 16;
  const int wm = wave / WGN, wn = wave % WGN;

  d4 acc[TI][TJ];
#pragma unroll
  for (int a = 0; a < TI; ++a)
#pragma unroll
    for (int b = 0; b < TJ; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};

  const bool a_kcontig = (G.a_cs == 1);
  const bool b_ncontig = (G.b_cs == 1);
  constexpr int A_PER = BM * BK / NT;
  constexpr int B_PER = BN * BK / NT;
  double ra[PF][A_PER], rb[PF][B_PER];  // panel p lives in slot p % PF
  // LDS element strides (k, i) of A and (k, j) of B for the chosen layouts
  const int a_ks = a_kcontig ? 1 : BM + PAD, a_is = a_kcontig ? BK + 1 : 1;
  const int b_ks = b_ncontig ? BN + PAD : 1, b_js = b_ncontig ? 1 : BK + 1;

  // store the panel starting at K index k0, zero past the split's end ke
  auto store = [&](int buf, const double (&sa)[A_PER], const double (&sb)[B_PER], int k0, int ke) {
#pragma unroll
    for (int q = 0; q < A_PER; ++q) {
      const int idx = threadIdx.x + NT * q;
      const int i = a_kcontig ? idx / BK : idx % BM;
      const int kk = a_kcontig ? idx % BK : idx / BM;
      As[buf][kk * a_ks + i * a_is] = k0 + kk < ke ? sa[q] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < B_PER; ++q) {
      const int idx = threadIdx.x + NT * q;
      const int j = b_ncontig ? idx % BN : idx / BK;
      const int kk = b_ncontig ? idx / BN : idx % BK;
      Bs[buf][kk * b_ks + j * b_js] = k0 + kk < ke ? sb[q] : 0.0;
    }
  };

  // split-K: blockIdx.y selects the K range [kb, ke) of this workgroup (empty
  // for the last splits of a group much shorter than the launch's longest K)
  const int kb = static_cast<int>(blockIdx.y) * G.kc;
  const int ke = min(G.k, kb + G.kc);
  const int np = ke > kb ? (ke - kb + BK - 1) / BK : 0;  // K panels
  auto mfma_panel = [&](int buf) {
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      const int kr = ks + (lane >> 4);
      double av[TI], bv[TJ];
#pragma unroll
      for (int a = 0; a < TI; ++a)
        av[a] = As[buf][kr * a_ks + (wm * WM + a * 16 + (lane & 15)) * a_is];
#pragma unroll
      for (int b = 0; b < TJ; ++b)
        bv[b] = Bs[buf][kr * b_ks + (wn * WN + b * 16 + (lane & 15)) * b_js];
#pragma unroll
      for (int a = 0; a < TI; ++a)
#pragma unroll
        for (int b = 0; b < TJ; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
  };
  // Panel p lives in register slot p % PF until it goes to LDS buffer p & 1.
  // The loads stay outside any branch (a load issued on one path only makes
  // the compiler wait for all loads at the join): the prologue loads panels
  // 0..PF, the main loop runs whole groups of PF panels, storing panel p + 1
  // and loading panel p + 1 + PF after each (clamped loads; panels past the
  // range store zeros that are never read), and the tail finishes the last
  // np % PF panels from the slots without loading.
  int cur = 0;
  if (G.k > 0) {  // uniform; k = 0 leaves the accumulators at zero
    load_a<BM, BK, A_PER, NT>(ra[0], A, G, m0, kb, a_kcontig);
    load_b<BN, BK, B_PER, NT>(rb[0], B, G, n0, kb, b_ncontig);
    store(0, ra[0], rb[0], kb, ke);
#pragma unroll
    for (int q = 1; q <= PF; ++q) {
      load_a<BM, BK, A_PER, NT>(ra[q % PF], A, G, m0, kb + q * BK, a_kcontig);
      load_b<BN, BK, B_PER, NT>(rb[q % PF], B, G, n0, kb + q * BK, b_ncontig);
    }
    __syncthreads();
    int p = 0;
    for (; p + PF <= np; p += PF) {
#pragma unroll
      for (int s = 0; s < PF; ++s) {
        const int slot = (s + 1) % PF;  // holds panel p + s + 1
        mfma_panel(cur);
        store(cur ^ 1, ra[slot], rb[slot], kb + (p + s + 1) * BK, ke);
        load_a<BM, BK, A_PER, NT>(ra[slot], A, G, m0, kb + (p + s + 1 + PF) * BK, a_kcontig);
        load_b<BN, BK, B_PER, NT>(rb[slot], B, G, n0, kb + (p + s + 1 + PF) * BK, b_ncontig);
        __syncthreads();
        cur ^= 1;
      }
    }
#pragma unroll
    for (int s = 0; s < PF - 1; ++s) {  // p is a multiple of PF: panel p + s + 1 in slot s + 1
      if (p + s < np) {
        mfma_panel(cur);
        if (p + s + 1 < np) store(cur ^ 1, ra[s + 1], rb[s + 1], kb + (p + s + 1) * BK, ke);
        __syncthreads();
        cur ^= 1;
      }
    }
  }

  // epilogue: for v_mfma_f64_16x16x4f64 lane l holds D[l/16 + 4*r][l%16] of
  // each 16x16 tile (rows interleaved by 4, measured on gfx950)
  if (bt.ws) {  // split-K partial, reduced by gemm_splitk_reduce
    double *__restrict__ P = bt.ws + G.ws_off + static_cast<int64_t>(blockIdx.y) * G.m * G.n;
#pragma unroll
    for (int a = 0; a < TI; ++a)
#pragma unroll
      for (int b = 0; b < TJ; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = m0 + wm * WM + a * 16 + (lane >> 4) + 4 * r;
          const int col = n0 + wn * WN + b * 16 + (lane & 15);
          if (row < G.m && col < G.n) P[static_cast<int64_t>(row) * G.n + col] = acc[a][b][r];
        }
    return;
  }
  double *__restrict__ C = G.c;
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

// C = alpha * sum_s partial_s + beta * C, one thread per output element; the
// splits are summed in a fixed order, so results do not depend on scheduling.
constexpr int kReduceThreads = 256;

__global__ void __launch_bounds__(kReduceThreads) gemm_splitk_reduce(Batch bt) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * kReduceThreads + threadIdx.x;
  if (e >= bt.elem_prefix[bt.ngroups]) return;
  int g = 0;
  while (g + 1 < bt.ngroups && bt.elem_prefix[g + 1] <= e) ++g;
  const GroupDev &G = bt.g[g];
  const int64_t local = e - bt.elem_prefix[g];
  const int64_t row = local / G.n, col = local - row * G.n;
  const double *p = bt.ws + G.ws_off + local;
  double sum = 0.0;
  const int64_t stride = static_cast<int64_t>(G.m) * G.n;
  for (int s = 0; s < G.splits; ++s) sum += p[s * stride];
  double *c = G.c + row * G.c_ld + col;
  const double v = bt.alpha * sum;
  *c = (bt.beta == 0.0) ? v : v + bt.beta * (*c);
}

// Wide products (m <= 128 rows, k <= 128, n long): C is a few rows over a very
// long n, e.g. a site tensor applied to a whole environment tensor. B is read
// once and C written once, with 2k flop per output element: both the HBM
// stream and the matrix cores are near their limits together, so the kernel
// keeps per-element overhead low by doing the arithmetic on FP64 MFMA
// (v_mfma_f64_16x16x4f64) with the whole of m in one workgroup. A workgroup
// stages all of A in LDS once (k-major, K padded to a multiple of 4, row
// stride MT*16 + 16 so the 4 k rows of one operand read fall in disjoint bank
// halves), then walks kWideSlabs slabs of 64 consecutive columns; in a slab
// wave w owns column tile w (16 columns) and all MT row tiles of 16 rows. Each
// wave reads its B operand straight from global memory (16 lanes read 16
// consecutive doubles of one K row), eight K steps' loads issued before their
// MFMAs. K rows past k read a clamped row of B against zero rows of A. C is
// written from the MFMA layout: each store covers four rows of 16
// consecutive doubles.
constexpr int kWideThreads = 256;
constexpr int kWideMaxM = 128;  // 8 row tiles
constexpr int kWideMaxK = 128;
constexpr int kWideCols = 64;   // columns per slab (4 waves x 16)
constexpr int kWideSlabs = 4;   // slabs per workgroup (A staged once for all)
constexpr int kWideLds = 64 * 1024;
constexpr int kWideStep = 8;    // K steps whose B loads are issued together

struct WideBatch {
  double alpha, beta;
  int32_t ngroups, pad;
  int32_t block_prefix[kMaxGroups + 1];
  GroupDev g[kMaxGroups];
};

constexpr int wide_lds_stride(int mt) { return mt * 16 + 16; }

template <int MT>
__global__ void __launch_bounds__(kWideThreads) gemm_wide_kernel(WideBatch bt) {
  constexpr int MR = MT * 16;
  constexpr int LS = wide_lds_stride(MT);
  extern __shared__ double As[];  // kpad x LS
  int gi = 0;
  while (gi + 1 < bt.ngroups && bt.block_prefix[gi + 1] <= static_cast<int>(blockIdx.x)) ++gi;
  const GroupDev &G = bt.g[gi];
  const int m = G.m, k = G.k;
  const int kpad = (k + 3) & ~3;
  for (int e = threadIdx.x; e < kpad * MR; e += kWideThreads) {
    const int kk = e / MR, i = e - kk * MR;
    As[kk * LS + i] = (i < m && kk < k)
                          ? G.a[static_cast<int64_t>(i) * G.a_rs + static_cast<int64_t>(kk) * G.a_cs]
                          : 0.0;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nsteps = kpad / 4;
  const double *arow0 = &As[(lane >> 4) * LS + (lane & 15)];
  const int64_t slab0 = static_cast<int64_t>(blockIdx.x - bt.block_prefix[gi]) * kWideSlabs;
  for (int sl = 0; sl < kWideSlabs; ++sl) {
    const int64_t j0 = (slab0 + sl) * kWideCols;
    if (j0 >= G.n) break;  // uniform
    const int64_t j = j0 + wave * 16 + (lane & 15);
    const double *__restrict__ bcol = G.b + min(j, static_cast<int64_t>(G.n) - 1) * G.b_cs;  // clamped column
    d4 acc[MT];
#pragma unroll
    for (int r = 0; r < MT; ++r) acc[r] = d4{0.0, 0.0, 0.0, 0.0};
    for (int s0 = 0; s0 < nsteps; s0 += kWideStep) {
      double bv[kWideStep];
#pragma unroll
      for (int u = 0; u < kWideStep; ++u)
        bv[u] = bcol[static_cast<int64_t>(min(4 * (s0 + u) + (lane >> 4), k - 1)) * G.b_rs];
#pragma unroll
      for (int u = 0; u < kWideStep; ++u) {
        if (s0 + u < nsteps) {  // uniform
          const double *arow = arow0 + 4 * (s0 + u) * LS;
#pragma unroll
          for (int r = 0; r < MT; ++r)
            acc[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(arow[16 * r], bv[u], acc[r], 0, 0, 0);
        }
      }
    }
    if (j < G.n) {
      // lane holds D[lane / 16 + 4 q][lane % 16] of each 16 x 16 tile
#pragma unroll
      for (int r = 0; r < MT; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int i = 16 * r + (lane >> 4) + 4 * q;
          if (i < m) {
            double *c = G.c + static_cast<int64_t>(i) * G.c_ld + j;
            const double v = bt.alpha * acc[r][q];
            *c = (bt.beta == 0.0) ? v : v + bt.beta * (*c);
          }
        }
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
  // split-K: a group with fewer splits than the launch leaves the rest idle
  // (whole workgroup, before any barrier)
  if (static_cast<int>(blockIdx.y) >= G.splits) return;
  const int t = tile - bt.tile_prefix[gi];
  const int tm = t / G.tiles_n;
  const int tn = t - tm * G.tiles_n;
  const int m0 = tm * BM;
  const int n0 = tn * BN;
  const double *__restrict__ A = G.a;
  const double *__restrict__ B = G.b;

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

  double *__restrict__ C = G.c;
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

constexpr int kBK = 16;

// Tile shape for one launch. Measured on MI355X (FP64, square m=n=k): 64x64
// reaches 42-47 TFLOP/s from 2048 up for every operand layout, 128x128 (one
// workgroup per CU at 398 registers) 30-39 TFLOP/s and far less on thin or
// small shapes, so 64x64 is the default; FS_GEMM_TILE=128 forces the other.
// 64 x 64 tiles by default; 32 x 32 when a launch would not cover the chip's
// 256 CUs with 64 x 64 tiles (small products, Gram partials), which
// quadruples the workgroups at the same total traffic through L2.
// FS_GEMM_TILE=32/64/128 forces a size (tuning only). Panel tiles spanning a
// thin side of up to 128 (64 x 128, the large operand streamed once) measured
// 20-45% slower than 32 x 32 tiles with four K panels in flight on the
// range-finder products (m x m times m x l, l = 20..69).
constexpr int kCUs = 256;

struct TileShape {
  int bm, bn;
};

// Static LDS of one workgroup of gemm_f64_kernel<BM, BN, kBK> (double
// buffered A and B panels, padded for either source layout).
constexpr int tile_lds_bytes(int bm, int bn) {
  const int a = (kBK * (bm + 16) > bm * (kBK + 1)) ? kBK * (bm + 16) : bm * (kBK + 1);
  const int b = (kBK * (bn + 16) > bn * (kBK + 1)) ? kBK * (bn + 16) : bn * (kBK + 1);
  return 2 * (a + b) * 8;
}

// Split-K: a launch whose output tiles leave most workgroup slots of the chip
// empty while its K is long (projected corners: chi x chi outputs over chi D^2;
// range-finder products: l x n outputs over m) runs several workgroups per
// tile over disjoint K ranges of at least kSplitMinK; partial tiles go to a
// workspace and a second kernel sums them in split order. Each workgroup keeps
// a few K panels in flight, so on such launches the bytes in flight, and with
// them the bandwidth reached, grow with the number of resident workgroups: the
// splits fill the resident slots (slots_per_cu workgroups per CU, set by the
// tile's LDS footprint out of 160 KB, at most 6). The 32 x 32 instance takes
// 124 registers, so 4 of its workgroups are resident per CU, not the 6 its LDS
// allows; shares sized for half the LDS-bound slots were 2 % slower
// (FS_GEMM_SPLITK_FILL=50), and holding the kernel to 94 registers for 5 per
// CU 3 % slower.
//
// The launch's K panels (tiles x ceil(k / kBK) summed over its groups) are
// dealt in shares of panels / slots: a group gets one split per share its K
// holds, so a group with a long K gets more splits than one with a short K
// and every workgroup runs about one share. With one split count for the
// whole launch, the tiles of the longest-K group would run longest while the
// short groups' workgroups idle their slots (the range-finder products hold
// sectors of 164 to 1412 rows and columns in one launch).
// FS_GEMM_SPLITK=0 disables split-K, FS_GEMM_SPLITK_BALANCE=0 gives every
// group the launch's uniform split count, FS_GEMM_SPLITK_FILL=P sizes the
// shares for P percent of the resident slots (tuning only).
constexpr int kSplitMinK = 256;
constexpr int kMaxSplits = 16;

inline int slots_per_cu(const TileShape &t) {
  if (t.bm == 128) return 1;  // register bound
  return std::max(1, std::min(6, (160 * 1024) / tile_lds_bytes(t.bm, t.bn)));
}

inline bool env_off(const char *name) {
  const char *e = std::getenv(name);
  return e && std::atoi(e) == 0;
}

// Split-K plan of one launch: sets kc, splits and ws_off of each group and the
// element prefix of the reduction; returns the launch's largest split count
// (grid.y, 1 without split-K) and the workspace size in doubles.
int plan_splits(Batch &bt, int64_t tiles, const TileShape &tile, int64_t *ws_doubles) {
  static const bool off = env_off("FS_GEMM_SPLITK");
  static const bool uniform = env_off("FS_GEMM_SPLITK_BALANCE");
  static const int64_t fill = [] {  // percent of the resident slots the shares target
    const char *e = std::getenv("FS_GEMM_SPLITK_FILL");
    const int v = e ? std::atoi(e) : 100;
    return static_cast<int64_t>(v >= 25 && v <= 400 ? v : 100);
  }();
  *ws_doubles = 0;
  const int ng = bt.ngroups;
  const int64_t slots = static_cast<int64_t>(kCUs) * slots_per_cu(tile);
  int64_t kmax = 0, panels = 0;
  for (int i = 0; i < ng; ++i) {
    const GroupDev &d = bt.g[i];
    kmax = std::max<int64_t>(kmax, d.k);
    panels += static_cast<int64_t>(bt.tile_prefix[i + 1] - bt.tile_prefix[i]) * fs::ceil_div(d.k, kBK);
  }
  if (off || tiles * 2 > slots || kmax < 2 * kSplitMinK) return 1;
  // uniform rule: one count for every group, by free slots and by the longest K
  const int64_t uni = std::max<int64_t>(1, std::min<int64_t>({slots / tiles, kmax / kSplitMinK, kMaxSplits}));
  const int64_t share = std::max<int64_t>(fs::ceil_div(panels * 100, slots * fill), kSplitMinK / kBK);
  int grid_y = 1;
  int64_t ws = 0, elems = 0;
  bt.elem_prefix[0] = 0;
  for (int i = 0; i < ng; ++i) {
    GroupDev &d = bt.g[i];
    const int64_t kp = fs::ceil_div(d.k, kBK);  // K panels of the group
    int64_t want = uniform ? uni : std::min<int64_t>(fs::ceil_div(kp, share), kMaxSplits);
    want = std::max<int64_t>(1, std::min<int64_t>(want, std::max<int64_t>(kp, 1)));
    d.kc = static_cast<int32_t>(std::max<int64_t>(fs::ceil_div(kp, want), 1) * kBK);
    d.splits = d.k > 0 ? static_cast<int32_t>(fs::ceil_div(d.k, d.kc)) : 1;
    d.ws_off = ws;
    ws += static_cast<int64_t>(d.splits) * d.m * d.n;
    elems += static_cast<int64_t>(d.m) * d.n;
    bt.elem_prefix[i + 1] = elems;
    grid_y = std::max(grid_y, static_cast<int>(d.splits));
  }
  if (grid_y == 1) {  // nothing split after all: plain launch
    for (int i = 0; i < ng; ++i) {
      bt.g[i].kc = static_cast<int32_t>(fs::ceil_div(bt.g[i].k, kBK) * kBK);
      bt.g[i].ws_off = 0;
    }
    return 1;
  }
  *ws_doubles = ws;
  return grid_y;
}

// Split-K workspaces: one per (device, stream), grow-only, process-wide, so
// host threads that come and go (serving slots) reuse them instead of each
// leaving one behind. A launch holds the pool's lock from the workspace lookup
// until both of its kernels (partial tiles, ordered reduction) are enqueued:
// threads sharing a stream then enqueue whole split-K pairs one after the
// other, and launches on other streams use other workspaces. The pool lives
// until the process ends (at most one buffer per stream that ever ran a
// split-K launch; torch hands out streams from a fixed pool).
struct SplitkWs {
  int device = -1;
  hipStream_t stream = nullptr;
  double *ptr = nullptr;
  int64_t size = 0;
};

struct SplitkPool {
  std::mutex mu;
  std::vector<SplitkWs> ws;
  std::vector<double *> retired;  // outgrown buffers, never freed
};

SplitkPool &splitk_pool() {
  static SplitkPool *pool = new SplitkPool;  // never destroyed: no HIP calls at exit
  return *pool;
}

// Caller holds splitk_pool().mu.
double *splitk_workspace_locked(int64_t n_doubles, hipStream_t stream) {
  SplitkPool &pool = splitk_pool();
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0) return nullptr;
  SplitkWs *found = nullptr;
  for (SplitkWs &x : pool.ws)
    if (x.device == dev && x.stream == stream) found = &x;
  if (!found) {
    pool.ws.push_back(SplitkWs{dev, stream, nullptr, 0});
    found = &pool.ws.back();
  }
  SplitkWs &w = *found;
  if (w.size < n_doubles) {
    // The outgrown buffer is kept, not freed: a recorded graph may hold its
    // address and replay later. Sizes at least double, so what is kept stays
    // below the final size.
    const int64_t want = std::max<int64_t>({n_doubles, 2 * w.size, int64_t(1) << 20});
    if (w.ptr) pool.retired.push_back(w.ptr);
    w.ptr = nullptr;
    w.size = 0;
    if (hipMalloc(&w.ptr, static_cast<size_t>(want) * sizeof(double)) != hipSuccess) return nullptr;
    w.size = want;
  }
  return w.ptr;
}

// Allocation-free entry point (fs_grouped_gemm_f64_noalloc): the split-K
// partial tiles live in this file-scope device array instead of a pool
// buffer. The split rule keeps a launch's partial tiles within its resident
// workgroup slots plus its output tiles (at most 2304 partial 32 x 32 tiles,
// 18 MiB; 64 x 64 tiles over a long K can need more); a launch that would
// need more than the array is refused. Each device holds its own copy at its own address, looked up once
// per device with hipGetSymbolAddress; host code uses only that address.
constexpr int64_t kScratchDoubles = int64_t(1) << 21;  // 16 MiB per device
static __device__ __align__(256) double fs_gemm_scratch[kScratchDoubles];

struct ScratchTable {
  std::mutex mu;                       // guards lookup and serializes split-K pairs
  std::vector<double *> ptr;           // per device: address of fs_gemm_scratch, or null
  std::vector<unsigned char> looked;   // per device: lookup done
};

ScratchTable &scratch_table() {
  static ScratchTable *t = new ScratchTable;  // never destroyed: no HIP calls at exit
  return *t;
}

// Caller holds scratch_table().mu. Address of the current device's scratch
// array (looked up on the first call on that device, with it current).
int scratch_ptr_locked(double **out) {
  ScratchTable &t = scratch_table();
  int dev = 0;
  FS_HIP_CHECK(hipGetDevice(&dev));
  FS_CHECK_ARG(dev >= 0, "grouped gemm: no current device");
  if (static_cast<size_t>(dev) >= t.ptr.size()) {
    t.ptr.resize(static_cast<size_t>(dev) + 1, nullptr);
    t.looked.resize(static_cast<size_t>(dev) + 1, 0);
  }
  if (!t.looked[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(fs_gemm_scratch));
    if (e != hipSuccess || p == nullptr) {
      fs::set_error("grouped gemm: hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
      return FS_ERR_HIP;
    }
    t.ptr[dev] = static_cast<double *>(p);
    t.looked[dev] = 1;
  }
  *out = t.ptr[dev];
  return FS_OK;
}

inline TileShape pick_tile(int ng, const fs_gemm_group *g) {
  static const int forced = [] {
    const char *e = std::getenv("FS_GEMM_TILE");
    return e ? std::atoi(e) : 0;
  }();
  if (forced == 32 || forced == 64 || forced == 128) return {forced, forced};
  int64_t tiles64 = 0;
  for (int i = 0; i < ng; ++i)
    if (g[i].m > 0 && g[i].n > 0) tiles64 += fs::ceil_div(g[i].m, 64) * fs::ceil_div(g[i].n, 64);
  const int t = tiles64 < kCUs ? 32 : 64;
  return {t, t};
}

// K panels in flight per workgroup (gemm_f64_kernel PF) for the 32 x 32
// tiles: their launches hold few workgroups per CU over long K ranges
// (range-finder products, Gram partials); four in flight measured 15-20%
// faster than one on the range-finder products (two and eight in between).
// The 64 x 64 tiles of the large square products keep one (register budget
// of 4 waves per SIMD).
constexpr int kSmallPrefetch = 4;

// Launch one tiled batch with the kernel instance of its tile shape.
void launch_tiles(const TileShape &t, dim3 grid, hipStream_t stream, const Batch &bt) {
  if (t.bm == 128)
    hipLaunchKernelGGL((gemm_f64_kernel<128, 128, kBK>), grid, dim3(kThreads), 0, stream, bt);
  else if (t.bm == 32)
    hipLaunchKernelGGL((gemm_f64_kernel<32, 32, kBK, 2, 2, kSmallPrefetch>), grid, dim3(kThreads), 0, stream, bt);
  else
    hipLaunchKernelGGL((gemm_f64_kernel<64, 64, kBK>), grid, dim3(kThreads), 0, stream, bt);
}

thread_local int last_dispatches = 0;

// Wide-kernel instance of a group (row tiles MT = 1..8 as index MT - 1), -1
// for the tile kernel: few rows and a short K over a long n, with B
// row-major (the 16 lanes of a K row read 16 consecutive columns) and A
// fitting the LDS. FS_GEMM_WIDE=0 disables it (tuning only).
constexpr int kNumWide = 8;

inline int wide_lds_bytes(int mt, int64_t k) { return static_cast<int>(((k + 3) & ~3) * wide_lds_stride(mt) * 8); }

inline int wide_instance(const fs_gemm_group &q) {
  static const bool off = [] {
    const char *e = std::getenv("FS_GEMM_WIDE");
    return e && std::atoi(e) == 0;
  }();
  if (off || q.m <= 0 || q.m > kWideMaxM || q.k <= 0 || q.k > kWideMaxK || q.b_cs != 1) return -1;
  // at least 16 slabs of columns (the tile kernel takes shorter products)
  if (q.n < 16 * kWideCols) return -1;
  const int mt = static_cast<int>((q.m + 15) / 16);
  if (wide_lds_bytes(mt, q.k) > kWideLds) return -1;
  return mt - 1;
}

template <int MT>
void launch_wide_mt(dim3 grid, int lds, hipStream_t stream, const WideBatch &wb) {
  hipLaunchKernelGGL((gemm_wide_kernel<MT>), grid, dim3(kWideThreads), lds, stream, wb);
}

void launch_wide(int w, dim3 grid, int lds, hipStream_t stream, const WideBatch &wb) {
  switch (w) {
    case 0: launch_wide_mt<1>(grid, lds, stream, wb); break;
    case 1: launch_wide_mt<2>(grid, lds, stream, wb); break;
    case 2: launch_wide_mt<3>(grid, lds, stream, wb); break;
    case 3: launch_wide_mt<4>(grid, lds, stream, wb); break;
    case 4: launch_wide_mt<5>(grid, lds, stream, wb); break;
    case 5: launch_wide_mt<6>(grid, lds, stream, wb); break;
    case 6: launch_wide_mt<7>(grid, lds, stream, wb); break;
    default: launch_wide_mt<8>(grid, lds, stream, wb); break;
  }
}

// Fill one device group descriptor (operand bases already resolved).
int fill_group(GroupDev &d, const fs_gemm_group &q, const double *a_base, const double *b_base,
               double *c_base) {
  FS_CHECK_ARG(q.m < (1ll << 31) && q.n < (1ll << 31) && q.k < (1ll << 31),
               "grouped gemm: dimension too large");
  FS_CHECK_ARG(a_base && b_base && c_base, "grouped gemm: null operand");
  FS_CHECK_ARG(q.a_off >= 0 && q.b_off >= 0 && q.c_off >= 0, "grouped gemm: negative offset");
  d.m = static_cast<int32_t>(q.m);
  d.n = static_cast<int32_t>(q.n);
  d.k = static_cast<int32_t>(q.k);
  d.a_rs = static_cast<int32_t>(q.a_rs);
  d.a_cs = static_cast<int32_t>(q.a_cs);
  d.b_rs = static_cast<int32_t>(q.b_rs);
  d.b_cs = static_cast<int32_t>(q.b_cs);
  d.c_ld = static_cast<int32_t>(q.c_ld);
  d.a = a_base + q.a_off;
  d.b = b_base + q.b_off;
  d.c = c_base + q.c_off;
  d.kc = static_cast<int32_t>(fs::ceil_div(q.k, kBK) * kBK);
  d.splits = 1;
  d.pad = 0;
  d.ws_off = 0;
  return FS_OK;
}

// A, B, C: per-group operand bases (stride 1 for one shared base, 0 entries
// otherwise unused); group offsets are relative to them. Wide groups (few
// rows over a long n) go to the streaming kernel of their row count, the
// others to the MFMA tile kernel; each launch takes at most kMaxGroups groups.
template <bool SPLIT>
int launch_grouped(int ngroups, const fs_gemm_group *groups, const double *const *A,
                   const double *const *B, double *const *C, int base_stride, double alpha,
                   double beta, void *stream, bool noalloc = false) {
  FS_CHECK_ARG(ngroups >= 0, "grouped gemm: negative group count");
  FS_CHECK_ARG(ngroups == 0 || groups, "grouped gemm: null groups");
  int dispatches = 0;
  std::vector<int> tiled, wide[kNumWide];
  for (int i = 0; i < ngroups; ++i) {
    const fs_gemm_group &q = groups[i];
    FS_CHECK_ARG(q.m >= 0 && q.n >= 0 && q.k >= 0, "grouped gemm: negative dimension");
    if (q.m == 0 || q.n == 0) continue;
    const int w = SPLIT ? -1 : wide_instance(q);
    (w >= 0 ? wide[w] : tiled).push_back(i);
  }
  for (int w = 0; w < kNumWide; ++w) {
    const int nc = kWideCols * kWideSlabs;  // columns per workgroup
    int lds = 0;
    for (size_t c0 = 0; c0 < wide[w].size(); c0 += kMaxGroups) {
      WideBatch wb;
      std::memset(&wb, 0, sizeof(wb));
      wb.alpha = alpha;
      wb.beta = beta;
      int ng = 0;
      int64_t blocks = 0;
      for (size_t t = c0; t < std::min(wide[w].size(), c0 + kMaxGroups); ++t) {
        const int i = wide[w][t];
        int st = fill_group(wb.g[ng], groups[i], A[i * base_stride], B[i * base_stride], C[i * base_stride]);
        if (st != FS_OK) return st;
        blocks += fs::ceil_div(groups[i].n, nc);
        FS_CHECK_ARG(blocks < (1ll << 31), "grouped gemm: too many blocks");
        lds = std::max(lds, wide_lds_bytes(w + 1, groups[i].k));
        wb.block_prefix[++ng] = static_cast<int32_t>(blocks);
      }
      wb.ngroups = ng;
      launch_wide(w, dim3(static_cast<unsigned>(blocks)), lds, fs::as_stream(stream), wb);
      FS_LAUNCH_CHECK("gemm_wide_kernel");
      ++dispatches;
    }
  }
  for (size_t c0 = 0; c0 < tiled.size(); c0 += kMaxGroups) {
    const size_t c1 = std::min(tiled.size(), c0 + kMaxGroups);
    std::vector<fs_gemm_group> sel;
    for (size_t t = c0; t < c1; ++t) sel.push_back(groups[tiled[t]]);
    const TileShape tile = SPLIT ? TileShape{64, 64} : pick_tile(static_cast<int>(sel.size()), sel.data());
    std::unique_lock<std::mutex> ws_lock;  // held from workspace lookup to the reduction launch
    Batch bt;
    std::memset(&bt, 0, sizeof(bt));
    bt.alpha = alpha;
    bt.beta = beta;
    int ng = 0;
    int64_t tiles = 0;
    bt.tile_prefix[0] = 0;
    for (size_t t = c0; t < c1; ++t) {
      const int i = tiled[t];
      const fs_gemm_group &q = groups[i];
      GroupDev &d = bt.g[ng];
      int st = fill_group(d, q, A[i * base_stride], B[i * base_stride], C[i * base_stride]);
      if (st != FS_OK) return st;
      d.tiles_n = static_cast<int32_t>(fs::ceil_div(q.n, tile.bn));
      tiles += fs::ceil_div(q.m, tile.bm) * d.tiles_n;
      FS_CHECK_ARG(tiles < (1ll << 31), "grouped gemm: too many tiles");
      bt.tile_prefix[ng + 1] = static_cast<int32_t>(tiles);
      ++ng;
    }
    bt.ngroups = ng;
    bt.xcd_remap = tiles >= 64 ? 1 : 0;
    int splits = 1;
    if (!SPLIT) {
      int64_t ws_doubles = 0;
      splits = plan_splits(bt, tiles, tile, &ws_doubles);
      if (splits > 1) {
        if (noalloc) {
          FS_CHECK_ARG(ws_doubles <= kScratchDoubles,
                       "grouped gemm (no-alloc): split-K needs %lld doubles, beyond the static "
                       "scratch of %lld",
                       static_cast<long long>(ws_doubles), static_cast<long long>(kScratchDoubles));
          ws_lock = std::unique_lock<std::mutex>(scratch_table().mu);
          const int st = scratch_ptr_locked(&bt.ws);
          if (st != FS_OK) return st;
        } else {
          ws_lock = std::unique_lock<std::mutex>(splitk_pool().mu);
          bt.ws = splitk_workspace_locked(ws_doubles, fs::as_stream(stream));
          if (bt.ws == nullptr) {
            fs::set_error("grouped gemm: split-K workspace allocation failed");
            return FS_ERR_HIP;
          }
        }
      }
    }
    const dim3 grid(static_cast<unsigned>(tiles), static_cast<unsigned>(splits));
    if (SPLIT)
      hipLaunchKernelGGL((gemm_split_kernel<64, 64, kBK>), grid, dim3(kThreads), 0,
                         fs::as_stream(stream), bt);
    else
      launch_tiles(tile, grid, fs::as_stream(stream), bt);
    FS_LAUNCH_CHECK(SPLIT ? "gemm_split_kernel" : "gemm_f64_kernel");
    ++dispatches;
    if (splits > 1) {
      const int64_t elems = bt.elem_prefix[ng];
      hipLaunchKernelGGL(gemm_splitk_reduce, dim3(static_cast<unsigned>(fs::ceil_div(elems, kReduceThreads))),
                         dim3(kReduceThreads), 0, fs::as_stream(stream), bt);
      FS_LAUNCH_CHECK("gemm_splitk_reduce");
      ++dispatches;
    }
    if (ws_lock.owns_lock()) ws_lock.unlock();
  }
  last_dispatches = dispatches;
  return FS_OK;
}

} // namespace

extern "C" {

int fs_grouped_gemm_f64(int ngroups, const fs_gemm_group *groups, const double *A,
                        const double *B, double *C, double alpha, double beta, void *stream) {
  return fs::guarded([&]() -> int {
    return launch_grouped<false>(ngroups, groups, &A, &B, &C, 0, alpha, beta, stream);
  });
}

int fs_grouped_gemm_f64_multi(int ngroups, const fs_gemm_group *groups, const double *const *A,
                              const double *const *B, double *const *C, double alpha, double beta,
                              void *stream) {
  return fs::guarded([&]() -> int {
    FS_CHECK_ARG(ngroups == 0 || (A && B && C), "grouped gemm: null base array");
    return launch_grouped<false>(ngroups, groups, A, B, C, 1, alpha, beta, stream);
  });
}

int fs_grouped_gemm_f64_noalloc(int ngroups, const fs_gemm_group *groups, const double *A,
                                const double *B, double *C, double alpha, double beta,
                                void *stream) {
  return fs::guarded([&]() -> int {
    return launch_grouped<false>(ngroups, groups, &A, &B, &C, 0, alpha, beta, stream, true);
  });
}

int fs_gemm_last_dispatches(void) { return last_dispatches; }

int fs_grouped_gemm_f64_split(int ngroups, const fs_gemm_group *groups, const double *A,
                              const double *B, double *C, double alpha, double beta,
                              void *stream) {
  return fs::guarded([&]() -> int {
    return launch_grouped<true>(ngroups, groups, &A, &B, &C, 0, alpha, beta, stream);
  });
}

} // extern "C"
