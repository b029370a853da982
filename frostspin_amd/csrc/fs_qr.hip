// Batched FP64 Householder QR, one workgroup per matrix.
//
// A (m x n) = Q R with q = min(m, n): Q (m x q) orthonormal columns, R (q x n)
// upper triangular, LAPACK geqrf/orgqr conventions (beta = -sign(alpha)*norm,
// so diagonal entries of R may be negative). The matrix is copied column-major
// into the work area; at step j every wavefront owns whole trailing columns
// (dot product with the reflector, then its rank-1 update), so a step needs only
// two workgroup barriers. Q is formed by backward accumulation of the
// reflectors. Used for the economic qr() of SymmetricTensor blocks
// (symmetric_tensor.py:655-671) and as the stable orthonormalisation step of
// the randomized truncated SVD.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "fs_hip.h"

namespace {

constexpr int kMaxMat = 24;
constexpr int kThreads = 512;
constexpr int kWaves = kThreads / 64;

struct QrDev {
  int32_t m, n;
  int32_t a_rs, a_cs;
  int64_t a_off, q_off, r_off, work_off;
};

struct QrBatch {
  const double *A;
  double *Q, *R, *work;
  int32_t nmat, want_r;
  QrDev d[kMaxMat];
};

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__global__ void __launch_bounds__(kThreads) householder_qr_kernel(QrBatch bt) {
  const QrDev D = bt.d[blockIdx.x];
  const int m = D.m, n = D.n;
  const int q = min(m, n);
  double *X = bt.work + D.work_off;                  // m x n column-major (reflectors below diag)
  double *tau = X + static_cast<int64_t>(m) * n;      // q
  double *Y = tau + q;                                // m x q column-major (Q accumulation)
  const double *A = bt.A + D.a_off;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  __shared__ double red[kWaves];
  __shared__ double s_beta, s_tau, s_v0;

  for (int64_t e = tid; e < static_cast<int64_t>(m) * n; e += kThreads) {
    const int c = static_cast<int>(e / m);
    const int r = static_cast<int>(e - static_cast<int64_t>(c) * m);
    X[e] = A[static_cast<int64_t>(r) * D.a_rs + static_cast<int64_t>(c) * D.a_cs];
  }
  __syncthreads();

  for (int j = 0; j < q; ++j) {
    double *xj = X + static_cast<int64_t>(j) * m;
    double part = 0.0;
    for (int i = j + 1 + tid; i < m; i += kThreads) part += xj[i] * xj[i];
    part = wave_sum(part);
    if (lane == 0) red[wave] = part;
    __syncthreads();
    if (tid == 0) {
      double sigma = 0.0;
      for (int w = 0; w < kWaves; ++w) sigma += red[w];
      const double alpha = xj[j];
      if (sigma == 0.0) {
        // already upper triangular in this column: H = I (LAPACK: tau = 0)
        s_tau = 0.0;
        s_beta = alpha;
        s_v0 = 1.0;
      } else {
        const double nrm = sqrt(alpha * alpha + sigma);
        const double beta = alpha >= 0.0 ? -nrm : nrm;
        s_tau = (beta - alpha) / beta;
        s_v0 = alpha - beta;  // v = [1, x(j+1:)/v0]
        s_beta = beta;
      }
    }
    __syncthreads();
    const double t = s_tau, v0 = s_v0;
    // scale the reflector so that v[0] = 1 (LAPACK dlarfg storage)
    if (t != 0.0)
      for (int i = j + 1 + tid; i < m; i += kThreads) xj[i] /= v0;
    __syncthreads();
    if (tid == 0) {
      xj[j] = s_beta;
      tau[j] = t;
    }
    // apply H = I - t v v^T to trailing columns, one wave per column
    if (t != 0.0) {
      for (int c = j + 1 + wave; c < n; c += kWaves) {
        double *xc = X + static_cast<int64_t>(c) * m;
        double d = (lane == 0) ? xc[j] : 0.0;
#pragma unroll 8
        for (int i = j + 1 + lane; i < m; i += 64) d += xj[i] * xc[i];
        d = wave_sum(d) * t;
        if (lane == 0) xc[j] -= d;
#pragma unroll 8
        for (int i = j + 1 + lane; i < m; i += 64) xc[i] -= d * xj[i];
      }
    }
    __syncthreads();
  }

  // Q = H_0 ... H_{q-1} [I; 0], backward accumulation into Y (m x q, column-major)
  for (int64_t e = tid; e < static_cast<int64_t>(m) * q; e += kThreads) {
    const int c = static_cast<int>(e / m);
    const int r = static_cast<int>(e - static_cast<int64_t>(c) * m);
    Y[e] = (r == c) ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int j = q - 1; j >= 0; --j) {
    const double t = tau[j];
    if (t == 0.0) continue;  // uniform across the workgroup
    const double *xj = X + static_cast<int64_t>(j) * m;
    for (int c = j + wave; c < q; c += kWaves) {
      double *yc = Y + static_cast<int64_t>(c) * m;
      double d = (lane == 0) ? yc[j] : 0.0;
#pragma unroll 8
      for (int i = j + 1 + lane; i < m; i += 64) d += xj[i] * yc[i];
      d = wave_sum(d) * t;
      if (lane == 0) yc[j] -= d;
#pragma unroll 8
      for (int i = j + 1 + lane; i < m; i += 64) yc[i] -= d * xj[i];
    }
    __syncthreads();
  }

  double *Qo = bt.Q + D.q_off;
  for (int64_t e = tid; e < static_cast<int64_t>(m) * q; e += kThreads) {
    const int r = static_cast<int>(e / q);
    const int c = static_cast<int>(e - static_cast<int64_t>(r) * q);
    Qo[e] = Y[static_cast<int64_t>(c) * m + r];
  }
  if (bt.want_r) {
    double *Ro = bt.R + D.r_off;
    for (int64_t e = tid; e < static_cast<int64_t>(q) * n; e += kThreads) {
      const int r = static_cast<int>(e / n);
      const int c = static_cast<int>(e - static_cast<int64_t>(r) * n);
      Ro[e] = (c >= r) ? X[static_cast<int64_t>(c) * m + r] : 0.0;
    }
  }
}

} // namespace

extern "C" {

int fs_batched_qr_f64(int nmat, const fs_qr_desc *descs, const double *A, double *Q, double *R,
                      double *work, void *stream) {
  return fs::guarded([&]() -> int {
    FS_CHECK_ARG(nmat >= 0, "batched qr: negative count");
    FS_CHECK_ARG(nmat == 0 || (descs && A && Q && work), "batched qr: null argument");
    int i0 = 0;
    while (i0 < nmat) {
      QrBatch bt;
      std::memset(&bt, 0, sizeof(bt));
      bt.A = A;
      bt.Q = Q;
      bt.R = R;
      bt.work = work;
      bt.want_r = R ? 1 : 0;
      int nb = 0;
      for (; i0 < nmat && nb < kMaxMat; ++i0) {
        const fs_qr_desc &s = descs[i0];
        FS_CHECK_ARG(s.m >= 1 && s.n >= 1, "batched qr: empty matrix");
        FS_CHECK_ARG(s.m < (1ll << 31) && s.n < (1ll << 31), "batched qr: matrix too large");
        QrDev &d = bt.d[nb++];
        d.m = static_cast<int32_t>(s.m);
        d.n = static_cast<int32_t>(s.n);
        d.a_rs = static_cast<int32_t>(s.a_rs);
        d.a_cs = static_cast<int32_t>(s.a_cs);
        d.a_off = s.a_off;
        d.q_off = s.q_off;
        d.r_off = s.r_off;
        d.work_off = s.work_off;
      }
      bt.nmat = nb;
      hipLaunchKernelGGL(householder_qr_kernel, dim3(nb), dim3(kThreads), 0, fs::as_stream(stream),
                         bt);
      FS_LAUNCH_CHECK("householder_qr_kernel");
    }
    return FS_OK;
  });
}

} // extern "C"
