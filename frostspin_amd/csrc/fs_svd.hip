// Batched FP64 SVD by one-sided (Hestenes) Jacobi, one workgroup per matrix.
//
// X (p x q, column-major, p = max(m,n), q = min(m,n)) is A or A^T. Pairs of
// columns are orthogonalised with the Rutishauser rotation; a round-robin
// tournament gives q/2 disjoint pairs per round so the wavefronts of the
// workgroup rotate different pairs concurrently (each pair: one wavefront,
// lanes striding over the column length). Sweeps repeat until no pair is
// rotated. When X and V (q x q) fit in LDS the whole iteration runs out of LDS;
// otherwise they live in the caller's work buffer (served by L2).
// One-sided Jacobi computes small singular values to high relative accuracy,
// which the s^-1/2 projector scaling of CTMRG relies on.
// Singular values are written sorted non-increasing like LAPACK gesdd
// (robust_svd, misc_tools/svd_tools.py:116-151); singular-vector signs are
// not canonical (neither are LAPACK's), tests compare gauge-invariant data.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "fs_hip.h"

namespace {

constexpr int kMaxMat = 24;
constexpr int kMaxQ = 1024;

struct SvdDev {
  int32_t m, n;
  int32_t a_rs, a_cs;
  int64_t a_off, u_off, s_off, vt_off, work_off;
};

struct SvdBatch {
  const double *A;
  double *U, *S, *VT, *work;
  int32_t nmat, max_sweeps;
  int32_t lds_doubles;  // >0: matrices with p*q+q*q <= lds_doubles run in LDS
  int32_t pad;
  SvdDev d[kMaxMat];
};

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum over an aligned group of 16 lanes (xor offsets < 16 never leave the group).
__device__ inline double group_sum16(double v) {
#pragma unroll
  for (int off = 8; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__global__ void __launch_bounds__(1024) jacobi_svd_kernel(SvdBatch bt) {
  extern __shared__ double lds[];
  __shared__ double sv[kMaxQ];
  __shared__ int rotated;
  const SvdDev D = bt.d[blockIdx.x];
  const int m = D.m, n = D.n;
  const bool tr = m < n;  // work on A^T when A is wide
  const int p = tr ? n : m;
  const int q = tr ? m : n;
  const int64_t pq = static_cast<int64_t>(p) * q;
  const bool in_lds = pq + static_cast<int64_t>(q) * q <= bt.lds_doubles;
  double *X = in_lds ? lds : bt.work + D.work_off;  // p x q, column-major
  double *V = X + pq;                                // q x q, column-major
  const double *A = bt.A + D.a_off;
  const int nthreads = blockDim.x;
  const int nwaves = nthreads >> 6;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int group = tid >> 4;  // 16-lane pair groups
  const int sub = tid & 15;
  const int ngroups = nthreads >> 4;

  for (int64_t e = tid; e < pq; e += nthreads) {
    const int col = static_cast<int>(e / p);
    const int row = static_cast<int>(e - static_cast<int64_t>(col) * p);
    const int i = tr ? col : row;  // row index in A
    const int j = tr ? row : col;  // col index in A
    X[e] = A[static_cast<int64_t>(i) * D.a_rs + static_cast<int64_t>(j) * D.a_cs];
  }
  for (int64_t e = tid; e < static_cast<int64_t>(q) * q; e += nthreads) {
    const int col = static_cast<int>(e / q);
    const int row = static_cast<int>(e - static_cast<int64_t>(col) * q);
    V[e] = (row == col) ? 1.0 : 0.0;
  }
  __shared__ double fro2_red[16];
  {
    double f = 0.0;
    for (int64_t e = tid; e < pq; e += nthreads) f += X[e] * X[e];
    f = wave_sum(f);
    if (lane == 0) fro2_red[wave] = f;
  }
  __syncthreads();
  double fro2 = 0.0;
  for (int w = 0; w < nwaves; ++w) fro2 += fro2_red[w];
  // columns whose squared norm is below (eps ||A||_F)^2 are numerical zeros of a
  // rank-deficient matrix: rotating them against each other only reshuffles
  // rounding noise and would keep the sweep loop from terminating
  const double negligible = fro2 * 4.930380657631324e-32;

  const int Q = (q + 1) & ~1;  // even number of players (one dummy if q odd)
  // rotate only if |cos angle(x_i, x_j)| > sqrt(p) * eps (LAPACK dgesvj criterion);
  // a smaller threshold sits below the rounding noise of the dot products and
  // never converges
  const double tol = sqrt(static_cast<double>(p)) * 2.220446049250313e-16;
  for (int sweep = 0; sweep < bt.max_sweeps && q > 1; ++sweep) {
    if (tid == 0) rotated = 0;
    __syncthreads();
    int any = 0;
    for (int r = 0; r < Q - 1; ++r) {
      // circle method: player Q-1 fixed, the others rotate. Each 16-lane group
      // of a wavefront owns one pair, so a wavefront rotates 4 pairs at once and
      // the dot-product reductions are 4 butterfly steps deep.
      for (int k = group; k < Q / 2; k += ngroups) {
        int i, j;
        if (k == 0) {
          i = Q - 1;
          j = r;
        } else {
          i = (r + k) % (Q - 1);
          j = (r - k + (Q - 1)) % (Q - 1);
        }
        if (i >= q || j >= q) continue;
        if (i > j) {
          const int s = i;
          i = j;
          j = s;
        }
        double *xi = X + static_cast<int64_t>(i) * p;
        double *xj = X + static_cast<int64_t>(j) * p;
        double a = 0.0, b = 0.0, g = 0.0;
#pragma unroll 4
        for (int t = sub; t < p; t += 16) {
          const double u = xi[t], v = xj[t];
          a += u * u;
          b += v * v;
          g += u * v;
        }
        // xor butterflies inside the 16-lane group: bit-identical sums on its lanes
        a = group_sum16(a);
        b = group_sum16(b);
        g = group_sum16(g);
        if (g != 0.0 && a > negligible && b > negligible && fabs(g) > tol * sqrt(a) * sqrt(b)) {
          const double zeta = (b - a) / (2.0 * g);
          const double t_ = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
          const double c = 1.0 / sqrt(1.0 + t_ * t_);
          const double s = c * t_;
#pragma unroll 4
          for (int t = sub; t < p; t += 16) {
            const double u = xi[t], v = xj[t];
            xi[t] = c * u - s * v;
            xj[t] = s * u + c * v;
          }
          double *vi = V + static_cast<int64_t>(i) * q;
          double *vj = V + static_cast<int64_t>(j) * q;
#pragma unroll 4
          for (int t = sub; t < q; t += 16) {
            const double u = vi[t], v = vj[t];
            vi[t] = c * u - s * v;
            vj[t] = s * u + c * v;
          }
          any = 1;
        }
      }
      __syncthreads();
    }
    if (any && sub == 0) atomicOr(&rotated, 1);
    __syncthreads();
    const int done = (rotated == 0);
    __syncthreads();
    if (done) break;
  }

  // singular values and their rank (descending, ties by index)
  for (int k = wave; k < q; k += nwaves) {
    const double *xk = X + static_cast<int64_t>(k) * p;
    double a = 0.0;
    for (int t = lane; t < p; t += 64) a += xk[t] * xk[t];
    a = wave_sum(a);
    if (lane == 0) sv[k] = sqrt(a);
  }
  __syncthreads();
  double *U = bt.U + D.u_off;
  double *VT = bt.VT + D.vt_off;
  for (int k = wave; k < q; k += nwaves) {
    const double sk = sv[k];
    int rank = 0;
    for (int l = lane; l < q; l += 64) {
      const double sl = sv[l];
      rank += (sl > sk || (sl == sk && l < k)) ? 1 : 0;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) rank += __shfl_xor(rank, off, 64);
    const double inv = sk > 0.0 ? 1.0 / sk : 0.0;
    const double *xk = X + static_cast<int64_t>(k) * p;
    const double *vk = V + static_cast<int64_t>(k) * q;
    if (!tr) {
      // A = U S V^T: U[:, rank] = x_k / s_k (m = p rows), VT[rank, :] = v_k
      for (int t = lane; t < p; t += 64) U[static_cast<int64_t>(t) * q + rank] = xk[t] * inv;
      for (int t = lane; t < q; t += 64) VT[static_cast<int64_t>(rank) * n + t] = vk[t];
    } else {
      // A^T = X: A = V S (x/s)^T: U[:, rank] = v_k (m = q rows), VT[rank, :] = x_k / s_k
      for (int t = lane; t < q; t += 64) U[static_cast<int64_t>(t) * q + rank] = vk[t];
      for (int t = lane; t < p; t += 64) VT[static_cast<int64_t>(rank) * n + t] = xk[t] * inv;
    }
    if (lane == 0) bt.S[D.s_off + rank] = sk;
  }
}

int lds_limit_doubles() {
  static int cached = -1;
  if (cached >= 0) return cached;
  int dev = 0, bytes = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&bytes, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
    bytes = 65536;
  // keep room for the static arrays (sv, rotated)
  const int usable = std::max(0, bytes - (kMaxQ * 8 + 256));
  if (hipFuncSetAttribute(reinterpret_cast<const void *>(jacobi_svd_kernel),
                          hipFuncAttributeMaxDynamicSharedMemorySize, usable) != hipSuccess) {
    (void)hipGetLastError();
    cached = 0;
    return cached;
  }
  cached = usable / 8;
  return cached;
}

} // namespace

extern "C" {

int fs_batched_svd_f64(int nmat, const fs_svd_desc *descs, const double *A, double *U,
                       double *S, double *VT, double *work, int max_sweeps, void *stream) {
  return fs::guarded([&]() -> int {
    FS_CHECK_ARG(nmat >= 0, "batched svd: negative count");
    FS_CHECK_ARG(nmat == 0 || (descs && A && U && S && VT), "batched svd: null argument");
    FS_CHECK_ARG(max_sweeps >= 1, "batched svd: max_sweeps must be >= 1");
    const int lds_cap = nmat ? lds_limit_doubles() : 0;
    int i0 = 0;
    while (i0 < nmat) {
      SvdBatch bt;
      std::memset(&bt, 0, sizeof(bt));
      bt.A = A;
      bt.U = U;
      bt.S = S;
      bt.VT = VT;
      bt.work = work;
      bt.max_sweeps = max_sweeps;
      int nb = 0;
      int64_t lds_need = 0;
      int qmax = 1;
      bool needs_work = false;
      for (; i0 < nmat && nb < kMaxMat; ++i0) {
        const fs_svd_desc &s = descs[i0];
        FS_CHECK_ARG(s.m >= 1 && s.n >= 1, "batched svd: empty matrix");
        FS_CHECK_ARG(s.m < (1ll << 31) && s.n < (1ll << 31), "batched svd: matrix too large");
        const int64_t q = std::min(s.m, s.n), p = std::max(s.m, s.n);
        FS_CHECK_ARG(q <= kMaxQ, "batched svd: min(m, n) must be <= 1024");
        const int64_t need = p * q + q * q;
        if (need <= lds_cap) lds_need = std::max(lds_need, need);
        else needs_work = true;
        qmax = static_cast<int>(std::max<int64_t>(qmax, q));
        SvdDev &d = bt.d[nb++];
        d.m = static_cast<int32_t>(s.m);
        d.n = static_cast<int32_t>(s.n);
        d.a_rs = static_cast<int32_t>(s.a_rs);
        d.a_cs = static_cast<int32_t>(s.a_cs);
        d.a_off = s.a_off;
        d.u_off = s.u_off;
        d.s_off = s.s_off;
        d.vt_off = s.vt_off;
        d.work_off = s.work_off;
      }
      FS_CHECK_ARG(!needs_work || work, "batched svd: work buffer required for large matrices");
      bt.nmat = nb;
      bt.lds_doubles = static_cast<int32_t>(lds_need);
      const int threads = qmax >= 64 ? 1024 : (qmax >= 32 ? 512 : 256);
      hipLaunchKernelGGL(jacobi_svd_kernel, dim3(nb), dim3(threads),
                         static_cast<size_t>(lds_need) * 8, fs::as_stream(stream), bt);
      FS_LAUNCH_CHECK("jacobi_svd_kernel");
    }
    return FS_OK;
  });
}

} // extern "C"
