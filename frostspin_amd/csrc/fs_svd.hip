// Batched FP64 SVD by one-sided (Hestenes) Jacobi, one workgroup per matrix.
//
// X (p x q, column-major, p = max(m,n), q = min(m,n)) is A or A^T. Pairs of
// columns are orthogonalised with the Rutishauser rotation; a round-robin
// tournament gives q/2 disjoint pairs per round, and each pair belongs to one
// 16-lane group of the workgroup (four pairs per wavefront), lanes striding
// over the column length. Sweeps repeat until no pair is rotated. When X and
// V (q x q) fit in LDS the whole iteration runs out of LDS; otherwise they
// live in the caller's work buffer (served by L2).
// Matrices in LDS with p <= 16 R take the register form: a group loads its two
// columns of X and of V once per round (R values each per lane), reduces the
// three dot products with DPP inside its row of 16 lanes, and rotates in
// registers, so a round costs one LDS read and at most one write per column.
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

// One DPP lane permutation of a double (two 32-bit moves) inside a row of 16
// lanes.
template <int CTRL>
__device__ inline double dpp_double(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// Sum over an aligned group of 16 lanes (one DPP row): row mirror (lane i with
// 15 - i), half-row mirror (i with 7 - i inside each half), then the quad
// swaps xor 2 and xor 1. Each step adds two values that its partner lane adds
// in the other order, so every lane of the group ends with the same bits.
__device__ inline double group_sum16(double v) {
  v += dpp_double<0x140>(v);  // row_mirror
  v += dpp_double<0x141>(v);  // row_half_mirror
  v += dpp_double<0x4E>(v);   // quad_perm [2, 3, 0, 1]
  v += dpp_double<0xB1>(v);   // quad_perm [1, 0, 3, 2]
  return v;
}

// Pair k of round r of the round-robin tournament over Q = Qm + 1 players (Q
// even): player Qm keeps its seat, the other Qm move one seat per round.
__device__ inline void tournament_pair(int r, int k, int Qm, int &i, int &j) {
  if (k == 0) {
    i = Qm;
    j = r;
  } else {
    i = r + k;
    if (i >= Qm) i -= Qm;
    j = r - k;
    if (j < 0) j += Qm;
  }
}

// Rotate the pair unless it is already orthogonal to working accuracy,
// |g| <= tol sqrt(a b) (LAPACK dgesvj criterion), or one column is a numerical
// zero. The test is squared unless the squared norms could overflow (large).
__device__ inline bool needs_rotation(double a, double b, double g, double negligible, double tol,
                                      bool large) {
  if (g == 0.0 || !(a > negligible) || !(b > negligible)) return false;
  return large ? fabs(g) > tol * sqrt(a) * sqrt(b) : g * g > (tol * tol) * (a * b);
}

// Rutishauser rotation for the Gram matrix [[a, g], [g, b]] of a column pair:
// t = sgn(zeta) / (|zeta| + sqrt(1 + zeta^2)), zeta = (b - a) / (2 g),
// c = 1 / sqrt(1 + t^2), s = c t. Unless the norms are large it is evaluated as
// t = sgn(d) e / (|d| + sqrt(d^2 + e^2)) with d = b - a, e = 2 g (sgn(0) = +1
// acting on |e|), one division and one square root fewer on the round's
// critical path.
__device__ inline void jacobi_rotation(double a, double b, double g, bool large, double &c,
                                       double &s) {
  double t;
  if (!large) {
    const double d = b - a, e = 2.0 * g;
    const double num = d > 0.0 ? e : (d < 0.0 ? -e : fabs(e));
    t = num / (fabs(d) + sqrt(d * d + e * e));
    c = rsqrt(1.0 + t * t);
  } else {
    const double zeta = (b - a) / (2.0 * g);
    t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
    c = 1.0 / sqrt(1.0 + t * t);
  }
  s = c * t;
}

// R > 0: register form for matrices in LDS with p <= 16 R (one pair per group
// needs q/2 <= groups, which the launch guarantees); R = 0: strided form only.
template <int R>
__global__ void __launch_bounds__(1024) jacobi_svd_kernel(SvdBatch bt) {
  extern __shared__ double lds[];
  __shared__ double sv[kMaxQ];
  __shared__ int rotated;
  __shared__ double fro2_red[16];
  const SvdDev D = bt.d[blockIdx.x];
  const int m = D.m, n = D.n;
  const bool tr = m < n;  // work on A^T when A is wide
  const int p = tr ? n : m;
  const int q = tr ? m : n;
  const int64_t pq = static_cast<int64_t>(p) * q;
  const bool in_lds = pq + static_cast<int64_t>(q) * q <= bt.lds_doubles;
  // X (p x q) and V (q x q), column-major, in LDS or in the work buffer. The
  // body is instantiated for each: through pointers derived from the LDS array
  // the compiler emits LDS instructions; through a pointer that may address
  // either it would emit flat accesses, which also wait on the memory counter.
  auto body = [&](double *X, double *V) __attribute__((always_inline)) {
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
    const int Qm = Q - 1;
    // rotate only if |cos angle(x_i, x_j)| > sqrt(p) * eps (LAPACK dgesvj criterion);
    // a smaller threshold sits below the rounding noise of the dot products and
    // never converges
    const double tol = sqrt(static_cast<double>(p)) * 2.220446049250313e-16;
    const bool large = fro2 > 1e150;  // squared column norms could overflow when squared again
    const bool regs = R > 0 && in_lds && p <= 16 * R && Q / 2 <= ngroups;
    for (int sweep = 0; sweep < bt.max_sweeps && q > 1; ++sweep) {
      if (tid == 0) rotated = 0;
      __syncthreads();
      int any = 0;
      for (int r = 0; r < Qm; ++r) {
        if (regs) {
          if constexpr (R > 0) {
            int i, j;
            tournament_pair(r, group, Qm, i, j);
            if (group < Q / 2 && i < q && j < q) {
              double *xi = X + static_cast<int64_t>(min(i, j)) * p;
              double *xj = X + static_cast<int64_t>(max(i, j)) * p;
              double *vi = V + static_cast<int64_t>(min(i, j)) * q;
              double *vj = V + static_cast<int64_t>(max(i, j)) * q;
              double u[R], v[R], vu[R], vv[R];
              // V is read with X (before the decision) so its LDS latency
              // overlaps the reductions
  #pragma unroll
              for (int e = 0; e < R; ++e) {
                const int t = sub + 16 * e;
                u[e] = t < p ? xi[t] : 0.0;
                v[e] = t < p ? xj[t] : 0.0;
                vu[e] = t < q ? vi[t] : 0.0;
                vv[e] = t < q ? vj[t] : 0.0;
              }
              double a = 0.0, b = 0.0, g = 0.0;
  #pragma unroll
              for (int e = 0; e < R; ++e) {
                a += u[e] * u[e];
                b += v[e] * v[e];
                g += u[e] * v[e];
              }
              a = group_sum16(a);
              b = group_sum16(b);
              g = group_sum16(g);
              if (needs_rotation(a, b, g, negligible, tol, large)) {
                double c, s;
                jacobi_rotation(a, b, g, large, c, s);
  #pragma unroll
                for (int e = 0; e < R; ++e) {
                  const int t = sub + 16 * e;
                  if (t < p) {
                    xi[t] = c * u[e] - s * v[e];
                    xj[t] = s * u[e] + c * v[e];
                  }
                  if (t < q) {
                    vi[t] = c * vu[e] - s * vv[e];
                    vj[t] = s * vu[e] + c * vv[e];
                  }
                }
                any = 1;
              }
            }
          }
        } else {
          // each 16-lane group owns one pair at a time
          for (int k = group; k < Q / 2; k += ngroups) {
            int i, j;
            tournament_pair(r, k, Qm, i, j);
            if (i >= q || j >= q) continue;
            double *xi = X + static_cast<int64_t>(min(i, j)) * p;
            double *xj = X + static_cast<int64_t>(max(i, j)) * p;
            double a = 0.0, b = 0.0, g = 0.0;
  #pragma unroll 4
            for (int t = sub; t < p; t += 16) {
              const double u = xi[t], v = xj[t];
              a += u * u;
              b += v * v;
              g += u * v;
            }
            a = group_sum16(a);
            b = group_sum16(b);
            g = group_sum16(g);
            if (needs_rotation(a, b, g, negligible, tol, large)) {
              double c, s;
              jacobi_rotation(a, b, g, large, c, s);
  #pragma unroll 4
              for (int t = sub; t < p; t += 16) {
                const double u = xi[t], v = xj[t];
                xi[t] = c * u - s * v;
                xj[t] = s * u + c * v;
              }
              double *vi = V + static_cast<int64_t>(min(i, j)) * q;
              double *vj = V + static_cast<int64_t>(max(i, j)) * q;
  #pragma unroll 4
              for (int t = sub; t < q; t += 16) {
                const double u = vi[t], v = vj[t];
                vi[t] = c * u - s * v;
                vj[t] = s * u + c * v;
              }
              any = 1;
            }
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
  };
  if (in_lds)
    body(lds, lds + pq);
  else
    body(bt.work + D.work_off, bt.work + D.work_off + pq);
}

// Instance for the register form of length R per lane (0: strided form only).
const void *kernel_for(int R) {
  switch (R) {
    case 2: return reinterpret_cast<const void *>(jacobi_svd_kernel<2>);
    case 4: return reinterpret_cast<const void *>(jacobi_svd_kernel<4>);
    case 6: return reinterpret_cast<const void *>(jacobi_svd_kernel<6>);
    case 8: return reinterpret_cast<const void *>(jacobi_svd_kernel<8>);
    default: return reinterpret_cast<const void *>(jacobi_svd_kernel<0>);
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
  for (const void *k : {kernel_for(0), kernel_for(2), kernel_for(4), kernel_for(6), kernel_for(8)}) {
    if (hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, usable) != hipSuccess) {
      (void)hipGetLastError();
      cached = 0;
      return cached;
    }
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
      int64_t preg = 0;  // longest column among matrices that take the register form
      bool needs_work = false;
      for (; i0 < nmat && nb < kMaxMat; ++i0) {
        const fs_svd_desc &s = descs[i0];
        FS_CHECK_ARG(s.m >= 1 && s.n >= 1, "batched svd: empty matrix");
        FS_CHECK_ARG(s.m < (1ll << 31) && s.n < (1ll << 31), "batched svd: matrix too large");
        const int64_t q = std::min(s.m, s.n), p = std::max(s.m, s.n);
        FS_CHECK_ARG(q <= kMaxQ, "batched svd: min(m, n) must be <= 1024");
        const int64_t need = p * q + q * q;
        if (need <= lds_cap) {
          lds_need = std::max(lds_need, need);
          if (p <= 128) preg = std::max(preg, p);
        } else {
          needs_work = true;
        }
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
      // q/2 pairs per round on 16-lane groups: 256 threads hold q < 32, 512 hold
      // q < 64 and 1024 hold q <= 128, the register form's limit
      const int threads = qmax >= 64 ? 1024 : (qmax >= 32 ? 512 : 256);
      const int R = preg == 0 ? 0 : (preg <= 32 ? 2 : (preg <= 64 ? 4 : (preg <= 96 ? 6 : 8)));
      void *args[] = {&bt};
      FS_HIP_CHECK(hipLaunchKernel(kernel_for(R), dim3(nb), dim3(threads), args,
                                   static_cast<size_t>(lds_need) * 8, fs::as_stream(stream)));
      FS_LAUNCH_CHECK("jacobi_svd_kernel");
    }
    return FS_OK;
  });
}

} // extern "C"
