// Cholesky factorisation of small SPD Gram matrices and the right triangular
// solve Y R^-1: the two non-GEMM steps of (shifted) CholeskyQR, which
// orthonormalises the tall-skinny range-finder bases of the randomized SVD with
// the bulk of its flops in grouped MFMA GEMMs (G = Y^T Y).
//
// Shifted CholeskyQR3 (Fukaya, Kannan, Nakatsukasa, Yamamoto, Yanagisawa,
// SIAM J. Sci. Comput. 42 (2020) A477): the first pass factors
// G + s I with s = shift_rel * trace(G) (shift_rel = 11 (m l + l (l+1)) u),
// which keeps the factorisation defined for cond(Y) up to ~u^-1; two more
// unshifted passes restore orthogonality to O(u).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "fs_hip.h"

namespace {

constexpr int kMaxMat = 24;
constexpr int kThreads = 512;

struct CholDev {
  int32_t l, pad;
  int64_t g_off, r_off;
  double shift_rel;
};

struct CholBatch {
  const double *G;
  double *R;
  int32_t *status;
  int32_t nmat, pad;
  CholDev d[kMaxMat];
};

// Right-looking upper Cholesky G = R^T R, one workgroup per matrix, in place in
// R (row-major l x l). status[b] = 0 on success, j+1 if pivot j is not positive.
__global__ void __launch_bounds__(kThreads) cholesky_kernel(CholBatch bt) {
  const CholDev D = bt.d[blockIdx.x];
  const int l = D.l;
  const double *G = bt.G + D.g_off;
  double *R = bt.R + D.r_off;
  const int tid = threadIdx.x;
  __shared__ double red[kThreads / 64];
  __shared__ int fail;
  // trace for the shift
  double tr = 0.0;
  for (int i = tid; i < l; i += kThreads) tr += G[static_cast<int64_t>(i) * l + i];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) tr += __shfl_xor(tr, off, 64);
  if ((tid & 63) == 0) red[tid >> 6] = tr;
  if (tid == 0) fail = 0;
  __syncthreads();
  double trace = 0.0;
  for (int w = 0; w < kThreads / 64; ++w) trace += red[w];
  const double shift = D.shift_rel * trace;
  // symmetrise from the upper triangle (G comes from a GEMM; its two halves can
  // differ in the last bit) and add the shift
  for (int64_t e = tid; e < static_cast<int64_t>(l) * l; e += kThreads) {
    const int i = static_cast<int>(e / l), c = static_cast<int>(e - static_cast<int64_t>(i) * l);
    const int a = min(i, c), b = max(i, c);
    R[e] = G[static_cast<int64_t>(a) * l + b] + (i == c ? shift : 0.0);
  }
  __syncthreads();
  for (int j = 0; j < l; ++j) {
    const double piv = R[static_cast<int64_t>(j) * l + j];
    if (!(piv > 0.0)) {  // also catches NaN
      if (tid == 0) fail = j + 1;
      break;  // piv is read by every thread after the same barrier: uniform
    }
    const double r = sqrt(piv);
    const double inv = 1.0 / r;
    __syncthreads();  // everyone has read the pivot before it is overwritten
    if (tid == 0) R[static_cast<int64_t>(j) * l + j] = r;
    for (int c = j + 1 + tid; c < l; c += kThreads) R[static_cast<int64_t>(j) * l + c] *= inv;
    __syncthreads();
    // trailing update of the upper triangle: R[i][c] -= R[j][i] R[j][c], j < i <= c
    const int w = l - j - 1;
    const int64_t cnt = static_cast<int64_t>(w) * (w + 1) / 2;
    for (int64_t e = tid; e < cnt; e += kThreads) {
      // map e to (ii, cc) with 0 <= ii <= cc < w (row-wise packing of the triangle)
      int ii = static_cast<int>((2.0 * w + 1.0 - sqrt((2.0 * w + 1.0) * (2.0 * w + 1.0) - 8.0 * e)) / 2.0);
      if (ii < 0) ii = 0;
      while (ii > 0 && static_cast<int64_t>(ii) * (2 * w - ii + 1) / 2 > e) --ii;
      while (static_cast<int64_t>(ii + 1) * (2 * w - ii) / 2 <= e) ++ii;
      const int cc = ii + static_cast<int>(e - static_cast<int64_t>(ii) * (2 * w - ii + 1) / 2);
      const int i = j + 1 + ii, c = j + 1 + cc;
      R[static_cast<int64_t>(i) * l + c] -=
          R[static_cast<int64_t>(j) * l + i] * R[static_cast<int64_t>(j) * l + c];
    }
    __syncthreads();
  }
  __syncthreads();
  // zero the strict lower triangle
  for (int64_t e = tid; e < static_cast<int64_t>(l) * l; e += kThreads) {
    const int i = static_cast<int>(e / l), c = static_cast<int>(e - static_cast<int64_t>(i) * l);
    if (c < i) R[e] = 0.0;
  }
  if (tid == 0) bt.status[blockIdx.x] = fail;
}

constexpr int kMaxSolve = 32;

struct SolveDev {
  int32_t m, l;
  int64_t y_off, y_rs, y_cs, r_off, o_off;
};

struct SolveBatch {
  const double *Y;
  const double *R;
  double *O;
  int32_t n, pad;
  int64_t block_prefix[kMaxSolve + 1];
  SolveDev d[kMaxSolve];
};

// O = Y R^-1 (R upper triangular, row-major l x l): one thread per row of Y,
// forward substitution x_j = (y_j - sum_{k<j} x_k R[k][j]) / R[j][j]. All lanes
// read the same R entries (broadcast loads); x lives in the output row.
__global__ void __launch_bounds__(256) trsm_rows_kernel(SolveBatch b) {
  const int64_t blk = blockIdx.x;
  int g = 0;
  while (g + 1 < b.n && b.block_prefix[g + 1] <= blk) ++g;
  const SolveDev D = b.d[g];
  const int64_t row = (blk - b.block_prefix[g]) * 256 + threadIdx.x;
  if (row >= D.m) return;
  const double *y = b.Y + D.y_off + row * D.y_rs;
  const double *R = b.R + D.r_off;
  double *x = b.O + D.o_off + row * D.l;
  for (int j = 0; j < D.l; ++j) {
    double acc = y[static_cast<int64_t>(j) * D.y_cs];
    for (int k = 0; k < j; ++k) acc -= x[k] * R[static_cast<int64_t>(k) * D.l + j];
    x[j] = acc / R[static_cast<int64_t>(j) * D.l + j];
  }
}

// Cholesky and triangular inverse of one small SPD Gram matrix per workgroup,
// entirely in LDS: A = sym(G) + shift I = R^T R, then R^-1 in place, so that
// the solve Y R^-1 of CholeskyQR becomes one grouped MFMA GEMM instead of a
// sequential substitution per row. Order l <= kCholInvMaxOrder.
constexpr int kCholInvMaxOrder = 128;
constexpr int kCholInvThreads = 256;

struct CholInvDev {
  int32_t l, n_part;
  int64_t g_off, g_stride, r_off, rinv_off;
  double shift_rel;
};

struct CholInvBatch {
  const double *G;
  double *R;  // may be null: R not written
  double *Rinv;
  int32_t *status;
  int32_t nmat, pad;
  CholInvDev d[kMaxMat];
};

// Register-tiled: a 16 x 16 grid of threads, thread (ty, tx) owns the 8 x 8
// elements (ty + 16p, tx + 16q) of the trailing matrix (later of X = R^-1).
// One barrier per step: the owners of the pivot row publish it unscaled in a
// double-buffered LDS row, then every thread applies the branch-free rank-1
// update to its registers (rows that must not change get a zero factor).
// Finished rows of R go to an LDS copy (row stride l_max + 1: the inverse reads
// its columns conflict-free); finished rows of X go straight to global memory.
//   Cholesky (right-looking): r_j = a_j / sqrt(a_jj);  a_ic -= a_ji a_jc / a_jj
//   inverse (R X = I, rows bottom-up): X_i = B_i / R_ii;  B_k -= R_ki X_i, k < i
constexpr int kCIGrid = 16;
constexpr int kCITile = kCholInvMaxOrder / kCIGrid;  // 8
constexpr int kCIRowStride = kCholInvMaxOrder + 1;
constexpr size_t kCholInvLds = static_cast<size_t>(kCholInvMaxOrder) * kCIRowStride * sizeof(double);

__global__ void __launch_bounds__(kCholInvThreads) cholinv_kernel(CholInvBatch bt) {
  extern __shared__ double Rl[];  // l x kCIRowStride: rows of R
  const CholInvDev D = bt.d[blockIdx.x];
  const int l = D.l;
  const double *G = bt.G + D.g_off;
  const int tx = threadIdx.x & (kCIGrid - 1);
  const int ty = threadIdx.x / kCIGrid;
  __shared__ double rowbuf[2][kCholInvMaxOrder];
  __shared__ double red[kCholInvThreads / 64];
  auto g_at = [&](int i, int k) {
    double v = 0.0;
    for (int p = 0; p < D.n_part; ++p) v += G[p * D.g_stride + static_cast<int64_t>(i) * l + k];
    return v;
  };
  double tr = 0.0;
  if (tx == ty) {
#pragma unroll
    for (int p = 0; p < kCITile; ++p) {
      const int i = ty + kCIGrid * p;
      if (i < l) tr += g_at(i, i);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) tr += __shfl_xor(tr, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = tr;
  __syncthreads();
  double trace = 0.0;
#pragma unroll
  for (int w = 0; w < kCholInvThreads / 64; ++w) trace += red[w];
  const double shift = D.shift_rel * trace;

  double a[kCITile][kCITile];
#pragma unroll
  for (int p = 0; p < kCITile; ++p) {
#pragma unroll
    for (int q = 0; q < kCITile; ++q) {
      const int i = ty + kCIGrid * p, c = tx + kCIGrid * q;
      // upper triangle of the symmetrised Gram matrix (+ shift); the lower
      // triangle of the registers is never read
      a[p][q] = (i < l && c < l && i <= c) ? g_at(i, c) + (i == c ? shift : 0.0) : 0.0;
    }
  }
  int fail = 0;
  for (int j = 0; j < l; ++j) {
    const int jp = j / kCIGrid, jt = j % kCIGrid, buf = j & 1;
    if (ty == jt) {
#pragma unroll
      for (int q = 0; q < kCITile; ++q) {
        double v = a[0][q];
#pragma unroll
        for (int p = 1; p < kCITile; ++p) v = p == jp ? a[p][q] : v;
        const int c = tx + kCIGrid * q;
        if (c < l) rowbuf[buf][c] = v;
      }
    }
    __syncthreads();
    const double piv = rowbuf[buf][j];
    if (!(piv > 0.0)) {  // uniform: every thread read the same value after the barrier
      fail = j + 1;
      break;
    }
    const double rinv = 1.0 / sqrt(piv);
    const double pinv = rinv * rinv;
    if (ty == 0) {
#pragma unroll
      for (int q = 0; q < kCITile; ++q) {
        const int c = tx + kCIGrid * q;
        if (c < l) Rl[j * kCIRowStride + c] = c < j ? 0.0 : (c == j ? piv * rinv : rowbuf[buf][c] * rinv);
      }
    }
    double rc[kCITile];
#pragma unroll
    for (int q = 0; q < kCITile; ++q) {
      const int c = tx + kCIGrid * q;
      rc[q] = (c > j && c < l) ? rowbuf[buf][c] : 0.0;
    }
#pragma unroll
    for (int p = 0; p < kCITile; ++p) {
      const int i = ty + kCIGrid * p;
      const double ri = (i > j && i < l) ? rowbuf[buf][i] * pinv : 0.0;
#pragma unroll
      for (int q = 0; q < kCITile; ++q) a[p][q] = fma(-ri, rc[q], a[p][q]);
    }
  }
  if (fail) {
    if (threadIdx.x == 0) bt.status[blockIdx.x] = fail;
    return;
  }
  __syncthreads();  // all rows of R are in LDS
  if (bt.R) {
    double *R = bt.R + D.r_off;
    for (int e = threadIdx.x; e < l * l; e += kCholInvThreads) {
      const int i = e / l, c = e - i * l;
      R[e] = Rl[i * kCIRowStride + c];
    }
  }
  double *X = bt.Rinv + D.rinv_off;
  double x[kCITile][kCITile];
#pragma unroll
  for (int p = 0; p < kCITile; ++p) {
#pragma unroll
    for (int q = 0; q < kCITile; ++q) x[p][q] = (ty + kCIGrid * p == tx + kCIGrid * q) ? 1.0 : 0.0;
  }
  for (int i = l - 1; i >= 0; --i) {
    const int ip = i / kCIGrid, it = i % kCIGrid, buf = i & 1;
    if (ty == it) {  // row i of X is final: scale, publish, store
      const double dinv = 1.0 / Rl[i * kCIRowStride + i];
#pragma unroll
      for (int q = 0; q < kCITile; ++q) {
        double v = x[0][q];
#pragma unroll
        for (int p = 1; p < kCITile; ++p) v = p == ip ? x[p][q] : v;
        const int c = tx + kCIGrid * q;
        if (c < l) {
          v = c >= i ? v * dinv : 0.0;
          rowbuf[buf][c] = v;
          X[i * l + c] = v;
        }
      }
    }
    __syncthreads();
    double rx[kCITile];
#pragma unroll
    for (int q = 0; q < kCITile; ++q) {
      const int c = tx + kCIGrid * q;
      rx[q] = (c >= i && c < l) ? rowbuf[buf][c] : 0.0;
    }
#pragma unroll
    for (int p = 0; p < kCITile; ++p) {
      const int k = ty + kCIGrid * p;
      const double ck = k < i ? Rl[k * kCIRowStride + i] : 0.0;
#pragma unroll
      for (int q = 0; q < kCITile; ++q) x[p][q] = fma(-ck, rx[q], x[p][q]);
    }
  }
  if (threadIdx.x == 0) bt.status[blockIdx.x] = 0;
}

// Blocked form (default): four columns of R per barrier. The owners publish
// four raw rows; every thread factors the 4 x 4 diagonal block redundantly,
// derives the four panel rows at its own rows and columns by forward
// substitution and applies the rank-4 update to its 8 x 4 register tile
// (512 threads as 16 x 32, thread (ty, tx) owns rows ty + 16p, columns
// tx + 32q). The inverse takes four rows of X per barrier the same way, with
// the owners of four columns of R publishing them. The matrix is padded with
// the identity up to a multiple of 4, so every step is branch-free.
constexpr int kC4Rows = 16, kC4Cols = 32, kC4Blk = 4;
constexpr int kC4P = kCholInvMaxOrder / kC4Rows;  // 8 rows per thread
constexpr int kC4Q = kCholInvMaxOrder / kC4Cols;  // 4 columns per thread
constexpr int kC4Threads = kC4Rows * kC4Cols;     // 512

// P x Q: the part of the 8 x 4 register tile that holds rows and columns
// below the padded order (16 P and 32 Q >= l rounded up to 4; the launch picks
// them from the batch's largest l); the rest would only carry identity padding.
template <int P, int Q>
__global__ void __launch_bounds__(kC4Threads) cholinv4_kernel(CholInvBatch bt) {
  const CholInvDev D = bt.d[blockIdx.x];
  const int l = D.l;
  const double *G = bt.G + D.g_off;
  const int tx = threadIdx.x % kC4Cols;
  const int ty = threadIdx.x / kC4Cols;
  __shared__ double rowbuf[2][kC4Blk][kCholInvMaxOrder];
  __shared__ double colbuf[2][kC4Blk][kCholInvMaxOrder];
  __shared__ double udiag[kCholInvMaxOrder / kC4Blk][kC4Blk * kC4Blk];
  __shared__ double red[kC4Threads / 64];
  __shared__ double gdiag[kCholInvMaxOrder];
  // upper triangle of G summed over its partials in partial order; all loads
  // of one partial are issued before any is used (one memory round trip per
  // partial, not one per element and partial)
  double a[P][Q];
#pragma unroll
  for (int p = 0; p < P; ++p)
#pragma unroll
    for (int q = 0; q < Q; ++q) a[p][q] = 0.0;
  for (int part = 0; part < D.n_part; ++part) {
    const double *Gp = G + part * D.g_stride;
    double v[P][Q];
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int i = ty + kC4Rows * p, c = tx + kC4Cols * q;
        v[p][q] = (i < l && c < l && i <= c) ? Gp[static_cast<int64_t>(i) * l + c] : 0.0;
      }
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
      for (int q = 0; q < Q; ++q) a[p][q] += v[p][q];
  }
  // trace for the shift: thread i adds G_ii (the diagonal reaches it through
  // LDS), then a reduction over the workgroup
#pragma unroll
  for (int p = 0; p < P; ++p)
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int i = ty + kC4Rows * p;
      if (i == tx + kC4Cols * q && i < l) gdiag[i] = a[p][q];
    }
  __syncthreads();
  double tr = threadIdx.x < l ? gdiag[threadIdx.x] : 0.0;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) tr += __shfl_xor(tr, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = tr;
  __syncthreads();
  double trace = 0.0;
#pragma unroll
  for (int w = 0; w < kC4Threads / 64; ++w) trace += red[w];
  const double shift = D.shift_rel * trace;
#pragma unroll
  for (int p = 0; p < P; ++p) {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int i = ty + kC4Rows * p, c = tx + kC4Cols * q;
      if (i < l && c < l)
        a[p][q] = (i == c) ? a[p][q] + shift : a[p][q];
      else
        a[p][q] = (i == c) ? 1.0 : 0.0;  // identity padding beyond l
    }
  }
  const int lpad = (l + kC4Blk - 1) / kC4Blk * kC4Blk;
  int fail = 0;
  for (int j0 = 0; j0 < lpad; j0 += kC4Blk) {
    const int buf = (j0 / kC4Blk) & 1;
    const int jp = j0 / kC4Rows, jt = j0 % kC4Rows;  // rows j0 .. j0+3: ty = jt .. jt+3, same p
    if (ty >= jt && ty < jt + kC4Blk) {
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        double v = a[0][q];
#pragma unroll
        for (int p = 1; p < P; ++p) v = (p == jp) ? a[p][q] : v;
        rowbuf[buf][ty - jt][tx + kC4Cols * q] = v;
      }
    }
    __syncthreads();
    // 4 x 4 upper Cholesky of the diagonal block (every thread)
    double U[kC4Blk][kC4Blk], uinv[kC4Blk];
#pragma unroll
    for (int r = 0; r < kC4Blk; ++r) {
      double d = rowbuf[buf][r][j0 + r];
#pragma unroll
      for (int k = 0; k < r; ++k) d = fma(-U[k][r], U[k][r], d);
      if (!(d > 0.0)) {  // uniform: same values in every thread
        fail = j0 + r + 1;
        break;
      }
      U[r][r] = sqrt(d);
      uinv[r] = 1.0 / U[r][r];
#pragma unroll
      for (int t = r + 1; t < kC4Blk; ++t) {
        double v = rowbuf[buf][r][j0 + t];
#pragma unroll
        for (int k = 0; k < r; ++k) v = fma(-U[k][r], U[k][t], v);
        U[r][t] = v * uinv[r];
      }
    }
    if (fail) break;
    // panel rows at own columns (Rc) and own rows (Ri): forward substitution
    double Rc[kC4Blk][Q], Ri[kC4Blk][P];
#pragma unroll
    for (int r = 0; r < kC4Blk; ++r) {
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int c = tx + kC4Cols * q;
        double v = rowbuf[buf][r][c];
#pragma unroll
        for (int k = 0; k < r; ++k) v = fma(-U[k][r], Rc[k][q], v);
        Rc[r][q] = (c >= j0 + kC4Blk) ? v * uinv[r] : 0.0;
      }
#pragma unroll
      for (int p = 0; p < P; ++p) {
        const int i = ty + kC4Rows * p;
        double v = rowbuf[buf][r][i];
#pragma unroll
        for (int k = 0; k < r; ++k) v = fma(-U[k][r], Ri[k][p], v);
        Ri[r][p] = (i >= j0 + kC4Blk) ? v * uinv[r] : 0.0;
      }
    }
    // rank-4 update (rows at or above the block get a zero factor)
#pragma unroll
    for (int p = 0; p < P; ++p) {
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        double upd = a[p][q];
#pragma unroll
        for (int r = 0; r < kC4Blk; ++r) upd = fma(-Ri[r][p], Rc[r][q], upd);
        a[p][q] = upd;
      }
    }
    // the owners of the block's rows store them final: U inside the block,
    // the panel row right of it, zero left of the diagonal
    if (ty >= jt && ty < jt + kC4Blk) {
      const int r = ty - jt;
      double urow[kC4Blk], prow[Q];
#pragma unroll
      for (int t = 0; t < kC4Blk; ++t) {
        double v = U[0][t];
#pragma unroll
        for (int k = 1; k < kC4Blk; ++k) v = (k == r) ? U[k][t] : v;
        urow[t] = (t >= r) ? v : 0.0;
      }
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        double v = Rc[0][q];
#pragma unroll
        for (int k = 1; k < kC4Blk; ++k) v = (k == r) ? Rc[k][q] : v;
        prow[q] = v;
      }
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int c = tx + kC4Cols * q;
        double fin = (c >= j0 + kC4Blk) ? prow[q] : 0.0;
#pragma unroll
        for (int t = 0; t < kC4Blk; ++t) fin = (c == j0 + t) ? urow[t] : fin;
#pragma unroll
        for (int p = 0; p < P; ++p)
          if (p == jp) a[p][q] = fin;
      }
    }
    if (threadIdx.x == 0) {
#pragma unroll
      for (int r = 0; r < kC4Blk; ++r)
#pragma unroll
        for (int t = 0; t < kC4Blk; ++t) udiag[j0 / kC4Blk][r * kC4Blk + t] = t >= r ? U[r][t] : 0.0;
    }
  }
  if (fail) {
    if (threadIdx.x == 0) bt.status[blockIdx.x] = fail;
    return;
  }
  if (bt.R) {
    double *R = bt.R + D.r_off;
#pragma unroll
    for (int p = 0; p < P; ++p) {
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int i = ty + kC4Rows * p, c = tx + kC4Cols * q;
        if (i < l && c < l) R[static_cast<int64_t>(i) * l + c] = (c >= i) ? a[p][q] : 0.0;
      }
    }
  }
  if (!bt.Rinv) {  // factor only
    if (threadIdx.x == 0) bt.status[blockIdx.x] = 0;
    return;
  }
  __syncthreads();  // udiag complete, row buffers free
  double x[P][Q];
#pragma unroll
  for (int p = 0; p < P; ++p) {
#pragma unroll
    for (int q = 0; q < Q; ++q) x[p][q] = (ty + kC4Rows * p == tx + kC4Cols * q) ? 1.0 : 0.0;
  }
  for (int i0 = lpad - kC4Blk; i0 >= 0; i0 -= kC4Blk) {
    const int buf = (i0 / kC4Blk) & 1;
    const int ip = i0 / kC4Rows, it = i0 % kC4Rows;
    const int cq = i0 / kC4Cols, ct = i0 % kC4Cols;  // columns i0 .. i0+3: tx = ct .. ct+3, same q
    if (ty >= it && ty < it + kC4Blk) {  // current rows i0 .. i0+3 of B
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        double v = x[0][q];
#pragma unroll
        for (int p = 1; p < P; ++p) v = (p == ip) ? x[p][q] : v;
        rowbuf[buf][ty - it][tx + kC4Cols * q] = v;
      }
    }
    if (tx >= ct && tx < ct + kC4Blk) {  // columns i0 .. i0+3 of R at rows k < i0
#pragma unroll
      for (int p = 0; p < P; ++p) {
        double v = a[p][0];
#pragma unroll
        for (int q = 1; q < Q; ++q) v = (q == cq) ? a[p][q] : v;
        const int k = ty + kC4Rows * p;
        colbuf[buf][tx - ct][k] = (k < i0) ? v : 0.0;
      }
    }
    __syncthreads();
    double U[kC4Blk][kC4Blk];
#pragma unroll
    for (int r = 0; r < kC4Blk; ++r)
#pragma unroll
      for (int t = 0; t < kC4Blk; ++t) U[r][t] = udiag[i0 / kC4Blk][r * kC4Blk + t];
    double Xc[kC4Blk][Q];
#pragma unroll
    for (int r = kC4Blk - 1; r >= 0; --r) {
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int c = tx + kC4Cols * q;
        double v = rowbuf[buf][r][c];
#pragma unroll
        for (int t = r + 1; t < kC4Blk; ++t) v = fma(-U[r][t], Xc[t][q], v);
        Xc[r][q] = (c >= i0) ? v / U[r][r] : 0.0;
      }
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int k = ty + kC4Rows * p;
#pragma unroll
      for (int r = 0; r < kC4Blk; ++r) {
        const double ck = colbuf[buf][r][k];
#pragma unroll
        for (int q = 0; q < Q; ++q) x[p][q] = fma(-ck, Xc[r][q], x[p][q]);
      }
    }
    if (ty >= it && ty < it + kC4Blk) {  // the block's rows of X are final
      const int r = ty - it;
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        double v = Xc[0][q];
#pragma unroll
        for (int k = 1; k < kC4Blk; ++k) v = (k == r) ? Xc[k][q] : v;
#pragma unroll
        for (int p = 0; p < P; ++p)
          if (p == ip) x[p][q] = v;
      }
    }
  }
  double *X = bt.Rinv + D.rinv_off;
#pragma unroll
  for (int p = 0; p < P; ++p) {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int i = ty + kC4Rows * p, c = tx + kC4Cols * q;
      if (i < l && c < l) X[static_cast<int64_t>(i) * l + c] = x[p][q];
    }
  }
  if (threadIdx.x == 0) bt.status[blockIdx.x] = 0;
}

// Instance whose register tile covers `rows16` blocks of 16 rows (and
// ceil(rows16 / 2) blocks of 32 columns).
const void *cholinv4_for(int rows16) {
  switch (rows16) {
    case 1: return reinterpret_cast<const void *>(cholinv4_kernel<1, 1>);
    case 2: return reinterpret_cast<const void *>(cholinv4_kernel<2, 1>);
    case 3: return reinterpret_cast<const void *>(cholinv4_kernel<3, 2>);
    case 4: return reinterpret_cast<const void *>(cholinv4_kernel<4, 2>);
    case 5: return reinterpret_cast<const void *>(cholinv4_kernel<5, 3>);
    case 6: return reinterpret_cast<const void *>(cholinv4_kernel<6, 3>);
    case 7: return reinterpret_cast<const void *>(cholinv4_kernel<7, 4>);
    default: return reinterpret_cast<const void *>(cholinv4_kernel<kC4P, kC4Q>);
  }
}

// Inverse X = R^-1 of one upper triangular l x l matrix (l <= 128) per
// workgroup, by block rows of 16 from the bottom: with B = I,
//   X_I = D_I B_I  (D_I = R_II^-1),  then  B_K -= R_KI X_I  for K < I.
// The diagonal-block inverses D_I are formed first (one thread per column of
// each, back substitution), so the sequential part is nb = ceil(l / 16) steps
// of two barriers each, instead of one barrier per row or four rows. 512
// threads as 16 x 32, thread (ty, tx) owns rows ty + 16p and columns tx + 32q
// of B / X in registers; block row I of B, the new X_I and the column block
// R_:I travel through double-buffered LDS. R is padded with the identity up to
// a multiple of 16. A matrix whose factorisation failed (status != 0) gets NaN.
constexpr int kTiBlk = 16;
constexpr int kTiRows = 16, kTiCols = 32;
constexpr int kTiThreads = kTiRows * kTiCols;        // 512
constexpr int kTiP = kCholInvMaxOrder / kTiRows;      // 8 rows per thread
constexpr int kTiQ = kCholInvMaxOrder / kTiCols;      // 4 columns per thread
constexpr int kTiNb = kCholInvMaxOrder / kTiBlk;      // 8 block rows

struct TriInvDev {
  int32_t l, pad;
  int64_t r_off, x_off;
};

struct TriInvBatch {
  const double *R;
  double *X;
  const int32_t *status;  // factorisation status per matrix (may be null)
  int32_t nmat, pad;
  TriInvDev d[kMaxMat];
};

// P x Q: the part of the 8 x 4 register tile below the order padded to 16
// (P = number of 16-row blocks, Q = ceil(P / 2); the launch picks them from
// the batch's largest l).
template <int P, int Q>
__global__ void __launch_bounds__(kTiThreads) triinv_kernel(TriInvBatch bt) {
  const TriInvDev D = bt.d[blockIdx.x];
  const int l = D.l;
  const double *R = bt.R + D.r_off;
  double *X = bt.X + D.x_off;
  const int tx = threadIdx.x % kTiCols;
  const int ty = threadIdx.x / kTiCols;
  if (bt.status && bt.status[blockIdx.x] != 0) {
    for (int e = threadIdx.x; e < l * l; e += kTiThreads) X[e] = __builtin_nan("");
    return;
  }
  const int nb = (l + kTiBlk - 1) / kTiBlk;
  __shared__ double diag[kTiNb][kTiBlk][kTiBlk + 1];             // R_II, then D_I
  __shared__ double rcol[2][kCholInvMaxOrder][kTiBlk + 1];       // R[k][16 I + t], k < 16 I
  __shared__ double brow[2][kTiBlk][kCholInvMaxOrder];           // block row I of B
  __shared__ double xrow[2][kTiBlk][kCholInvMaxOrder];           // X_I
  auto r_at = [&](int i, int k) -> double {
    if (i < l && k < l) return R[static_cast<int64_t>(i) * l + k];
    return i == k ? 1.0 : 0.0;
  };
  for (int e = threadIdx.x; e < nb * kTiBlk * kTiBlk; e += kTiThreads) {
    const int I = e / (kTiBlk * kTiBlk), r = (e / kTiBlk) % kTiBlk, t = e % kTiBlk;
    diag[I][r][t] = (t >= r) ? r_at(I * kTiBlk + r, I * kTiBlk + t) : 0.0;
  }
  __syncthreads();
  // D_I column c: R_II d = e_c by back substitution (thread (I, c))
  const int di = threadIdx.x / kTiBlk, dc = threadIdx.x % kTiBlk;
  double dcol[kTiBlk];
  if (di < nb) {
#pragma unroll
    for (int r = kTiBlk - 1; r >= 0; --r) {
      double v = (r == dc) ? 1.0 : 0.0;
#pragma unroll
      for (int t = r + 1; t < kTiBlk; ++t) v = fma(-diag[di][r][t], dcol[t], v);
      dcol[r] = v / diag[di][r][r];
    }
  }
  __syncthreads();  // every R_II read before D_I overwrites it
  if (di < nb) {
#pragma unroll
    for (int r = 0; r < kTiBlk; ++r) diag[di][r][dc] = dcol[r];
  }
  double x[P][Q];
#pragma unroll
  for (int p = 0; p < P; ++p)
#pragma unroll
    for (int q = 0; q < Q; ++q) x[p][q] = (ty + kTiRows * p == tx + kTiCols * q) ? 1.0 : 0.0;

  for (int I = nb - 1; I >= 0; --I) {
    const int buf = I & 1;
    // publish block row I of B (every thread owns one of its rows) and the
    // column block R_:I above the diagonal block
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      double v = x[0][q];
#pragma unroll
      for (int p = 1; p < P; ++p) v = (p == I) ? x[p][q] : v;
      brow[buf][ty][tx + kTiCols * q] = v;
    }
    for (int e = threadIdx.x; e < I * kTiBlk * kTiBlk; e += kTiThreads) {
      const int k = e / kTiBlk, t = e % kTiBlk;
      rcol[buf][k][t] = r_at(k, I * kTiBlk + t);
    }
    __syncthreads();
    // X_I = D_I B_I at own row ty of the block and own columns
    double xi[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) xi[q] = 0.0;
    for (int t = ty; t < kTiBlk; ++t) {
      const double dv = diag[I][ty][t];
#pragma unroll
      for (int q = 0; q < Q; ++q) xi[q] = fma(dv, brow[buf][t][tx + kTiCols * q], xi[q]);
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      xrow[buf][ty][tx + kTiCols * q] = xi[q];
#pragma unroll
      for (int p = 0; p < P; ++p)
        if (p == I) x[p][q] = xi[q];
    }
    __syncthreads();
    // rows of the blocks above: B_K -= R_KI X_I
#pragma unroll
    for (int p = 0; p < P; ++p) {
      if (p < I) {
        const int k = ty + kTiRows * p;
#pragma unroll 4
        for (int t = 0; t < kTiBlk; ++t) {
          const double rk = rcol[buf][k][t];
#pragma unroll
          for (int q = 0; q < Q; ++q) x[p][q] = fma(-rk, xrow[buf][t][tx + kTiCols * q], x[p][q]);
        }
      }
    }
  }
#pragma unroll
  for (int p = 0; p < P; ++p) {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int i = ty + kTiRows * p, c = tx + kTiCols * q;
      if (i < l && c < l) X[static_cast<int64_t>(i) * l + c] = x[p][q];
    }
  }
}

const void *triinv_for(int rows16) {
  switch (rows16) {
    case 1: return reinterpret_cast<const void *>(triinv_kernel<1, 1>);
    case 2: return reinterpret_cast<const void *>(triinv_kernel<2, 1>);
    case 3: return reinterpret_cast<const void *>(triinv_kernel<3, 2>);
    case 4: return reinterpret_cast<const void *>(triinv_kernel<4, 2>);
    case 5: return reinterpret_cast<const void *>(triinv_kernel<5, 3>);
    case 6: return reinterpret_cast<const void *>(triinv_kernel<6, 3>);
    case 7: return reinterpret_cast<const void *>(triinv_kernel<7, 4>);
    default: return reinterpret_cast<const void *>(triinv_kernel<kTiP, kTiQ>);
  }
}

} // namespace

extern "C" {

int fs_batched_cholinv_f64(int nmat, const fs_cholinv_desc *descs, const double *G, double *R,
                           double *Rinv, int32_t *status, void *stream) {
  return fs::guarded([&]() -> int {
    FS_CHECK_ARG(nmat >= 0, "batched cholinv: negative count");
    FS_CHECK_ARG(nmat == 0 || (descs && G && (Rinv || R) && status), "batched cholinv: null argument");
    for (int i = 0; i < nmat; ++i) {
      FS_CHECK_ARG(descs[i].l >= 1 && descs[i].l <= kCholInvMaxOrder,
                   "batched cholinv: order must be in [1, %d]", kCholInvMaxOrder);
      FS_CHECK_ARG(descs[i].shift_rel >= 0.0, "batched cholinv: negative shift");
      FS_CHECK_ARG(descs[i].n_part >= 1 && descs[i].n_part <= (1 << 16) && descs[i].g_stride >= 0,
                   "batched cholinv: bad partial count / stride");
    }
    static thread_local int attr_device = -1;
    int dev = 0;
    FS_HIP_CHECK(hipGetDevice(&dev));
    if (attr_device != dev) {
      FS_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(cholinv_kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       static_cast<int>(kCholInvLds)));
      attr_device = dev;
    }
    int i0 = 0;
    while (i0 < nmat) {
      CholInvBatch bt;
      std::memset(&bt, 0, sizeof(bt));
      bt.G = G;
      bt.R = R;
      bt.Rinv = Rinv;
      bt.status = status + i0;
      int nb = 0;
      for (; i0 < nmat && nb < kMaxMat; ++i0) {
        const fs_cholinv_desc &s = descs[i0];
        CholInvDev &d = bt.d[nb++];
        d.l = static_cast<int32_t>(s.l);
        d.n_part = static_cast<int32_t>(s.n_part);
        d.g_off = s.g_off;
        d.g_stride = s.g_stride;
        d.r_off = s.r_off;
        d.rinv_off = s.rinv_off;
        d.shift_rel = s.shift_rel;
      }
      bt.nmat = nb;
      static const int mode = [] {
        const char *e = std::getenv("FS_CHOLINV_KERNEL");  // tuning only: 1 = one column per step
        return e ? std::atoi(e) : 0;
      }();
      if (mode == 1 && Rinv) {
        hipLaunchKernelGGL(cholinv_kernel, dim3(nb), dim3(kCholInvThreads), kCholInvLds,
                           fs::as_stream(stream), bt);
        FS_LAUNCH_CHECK("cholinv_kernel");
      } else {
        int lmax = 1;
        for (int b = 0; b < nb; ++b) lmax = std::max<int>(lmax, bt.d[b].l);
        const int rows16 = (((lmax + kC4Blk - 1) / kC4Blk * kC4Blk) + kC4Rows - 1) / kC4Rows;  // P
        void *args[] = {&bt};
        FS_HIP_CHECK(hipLaunchKernel(cholinv4_for(rows16), dim3(nb), dim3(kC4Threads), args, 0,
                                     fs::as_stream(stream)));
        FS_LAUNCH_CHECK("cholinv4_kernel");
      }
    }
    return FS_OK;
  });
}

int fs_batched_cholesky_f64(int nmat, const fs_chol_desc *descs, const double *G, double *R,
                            int32_t *status, void *stream) {
  return fs::guarded([&]() -> int {
    FS_CHECK_ARG(nmat >= 0, "batched cholesky: negative count");
    FS_CHECK_ARG(nmat == 0 || (descs && G && R && status), "batched cholesky: null argument");
    int i0 = 0;
    while (i0 < nmat) {
      CholBatch bt;
      std::memset(&bt, 0, sizeof(bt));
      bt.G = G;
      bt.R = R;
      bt.status = status + i0;
      int nb = 0;
      for (; i0 < nmat && nb < kMaxMat; ++i0) {
        const fs_chol_desc &s = descs[i0];
        FS_CHECK_ARG(s.l >= 1 && s.l < (1 << 15), "batched cholesky: bad order");
        FS_CHECK_ARG(s.shift_rel >= 0.0, "batched cholesky: negative shift");
        CholDev &d = bt.d[nb++];
        d.l = static_cast<int32_t>(s.l);
        d.g_off = s.g_off;
        d.r_off = s.r_off;
        d.shift_rel = s.shift_rel;
      }
      bt.nmat = nb;
      hipLaunchKernelGGL(cholesky_kernel, dim3(nb), dim3(kThreads), 0, fs::as_stream(stream), bt);
      FS_LAUNCH_CHECK("cholesky_kernel");
    }
    return FS_OK;
  });
}

int fs_batched_trsm_right_upper_f64(int nmat, const fs_trsm_desc *descs, const double *Y,
                                    const double *R, double *O, void *stream) {
  return fs::guarded([&]() -> int {
    FS_CHECK_ARG(nmat >= 0, "batched trsm: negative count");
    FS_CHECK_ARG(nmat == 0 || (descs && Y && R && O), "batched trsm: null argument");
    int i0 = 0;
    while (i0 < nmat) {
      SolveBatch b;
      std::memset(&b, 0, sizeof(b));
      b.Y = Y;
      b.R = R;
      b.O = O;
      int nb = 0;
      int64_t blocks = 0;
      for (; i0 < nmat && nb < kMaxSolve; ++i0) {
        const fs_trsm_desc &s = descs[i0];
        FS_CHECK_ARG(s.m >= 0 && s.l >= 1, "batched trsm: bad shape");
        if (s.m == 0) continue;
        SolveDev &d = b.d[nb];
        d.m = static_cast<int32_t>(s.m);
        d.l = static_cast<int32_t>(s.l);
        d.y_off = s.y_off;
        d.y_rs = s.y_rs;
        d.y_cs = s.y_cs;
        d.r_off = s.r_off;
        d.o_off = s.o_off;
        b.block_prefix[nb] = blocks;
        blocks += fs::ceil_div(s.m, 256);
        ++nb;
      }
      if (nb == 0) continue;
      b.block_prefix[nb] = blocks;
      b.n = nb;
      hipLaunchKernelGGL(trsm_rows_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0,
                         fs::as_stream(stream), b);
      FS_LAUNCH_CHECK("trsm_rows_kernel");
    }
    return FS_OK;
  });
}

int fs_batched_triinv_f64(int nmat, const fs_triinv_desc *descs, const double *R, double *X,
                          const int32_t *status, void *stream) {
  return fs::guarded([&]() -> int {
    FS_CHECK_ARG(nmat >= 0, "batched triinv: negative count");
    FS_CHECK_ARG(nmat == 0 || (descs && R && X), "batched triinv: null argument");
    for (int i = 0; i < nmat; ++i)
      FS_CHECK_ARG(descs[i].l >= 1 && descs[i].l <= kCholInvMaxOrder && descs[i].r_off >= 0 &&
                       descs[i].x_off >= 0,
                   "batched triinv: order must be in [1, %d]", kCholInvMaxOrder);
    int i0 = 0;
    while (i0 < nmat) {
      TriInvBatch bt;
      std::memset(&bt, 0, sizeof(bt));
      bt.R = R;
      bt.X = X;
      bt.status = status ? status + i0 : nullptr;
      int nb = 0;
      for (; i0 < nmat && nb < kMaxMat; ++i0) {
        TriInvDev &d = bt.d[nb++];
        d.l = static_cast<int32_t>(descs[i0].l);
        d.r_off = descs[i0].r_off;
        d.x_off = descs[i0].x_off;
      }
      bt.nmat = nb;
      int lmax = 1;
      for (int b = 0; b < nb; ++b) lmax = std::max<int>(lmax, bt.d[b].l);
      void *args[] = {&bt};
      FS_HIP_CHECK(hipLaunchKernel(triinv_for((lmax + kTiBlk - 1) / kTiBlk), dim3(nb), dim3(kTiThreads), args,
                                   0, fs::as_stream(stream)));
      FS_LAUNCH_CHECK("triinv_kernel");
    }
    return FS_OK;
  });
}

} // extern "C"
