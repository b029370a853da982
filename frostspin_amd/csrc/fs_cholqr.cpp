// Multi-pass shifted CholeskyQR issued from the host side of the C ABI.
//
// One call orthonormalises a batch of tall matrices: the descriptors of every
// launch of every pass are built here once (the passes after the first read
// the packed output of the one before, so they share their layout) and the
// launches go back to back on one stream: split-K Gram product (grouped GEMM),
// factor-only Cholesky, blocked triangular inverse, Y R^-1 (grouped GEMM), and
// with want_r the running product of the passes' factors. The launches and
// their descriptors are the ones a caller issuing the four steps one by one
// would make, so the results are identical; only the host work per pass goes.
#include <algorithm>
#include <cfloat>
#include <vector>

#include "fs_common.h"

namespace {

constexpr int64_t kGramRowsPerPart = 192;  // rows of Y per split-K Gram partial
constexpr int64_t kGramMaxParts = 16;
constexpr int64_t kMaxOrder = 128;  // fs_batched_cholinv_f64 / fs_batched_triinv_f64 limit
constexpr int kMaxPasses = 3;

// Offsets (in doubles) of every per-matrix region of the work buffer.
struct Layout {
  std::vector<int64_t> nparts, g_off, l2_off, ml_off;
  int64_t g_base = 0, r_base = 0, x_base = 0, t_base[2] = {0, 0}, acc_base = 0, total = 0;
  int64_t l2_total = 0, ml_total = 0;
};

int make_layout(int nmat, const fs_cholqr_desc *d, int passes, bool want_r, Layout &L) {
  FS_CHECK_ARG(nmat >= 0, "cholqr: negative count");
  FS_CHECK_ARG(nmat == 0 || d, "cholqr: null descriptors");
  FS_CHECK_ARG(passes >= 1 && passes <= kMaxPasses, "cholqr: passes must be 1, 2 or 3");
  FS_CHECK_ARG(!want_r || passes >= 2, "cholqr: the triangular factor needs at least two passes");
  L.nparts.resize(nmat);
  L.g_off.resize(nmat);
  L.l2_off.resize(nmat);
  L.ml_off.resize(nmat);
  int64_t g = 0, l2 = 0, ml = 0;
  for (int i = 0; i < nmat; ++i) {
    const int64_t m = d[i].m, l = d[i].l;
    FS_CHECK_ARG(m >= 1 && l >= 1 && l <= kMaxOrder, "cholqr: matrix %d is %lld x %lld (need m >= 1, 1 <= l <= 128)",
                 i, static_cast<long long>(m), static_cast<long long>(l));
    const int64_t np = std::clamp<int64_t>((m + kGramRowsPerPart - 1) / kGramRowsPerPart, 1, kGramMaxParts);
    L.nparts[i] = np;
    L.g_off[i] = g;
    L.l2_off[i] = l2;
    L.ml_off[i] = ml;
    g += np * l * l;
    l2 += l * l;
    ml += m * l;
  }
  L.l2_total = l2;
  L.ml_total = ml;
  // G partials | R (one per pass with want_r) | X | T0 | T1 | running product
  L.g_base = 0;
  L.r_base = g;
  L.x_base = L.r_base + (want_r ? passes : 1) * l2;
  int64_t next = L.x_base + l2;
  for (int t = 0; t < std::min(passes - 1, 2); ++t) {
    L.t_base[t] = next;
    next += ml;
  }
  L.acc_base = next;
  if (want_r && passes >= 3) next += l2;
  L.total = next;
  return FS_OK;
}

} // namespace

extern "C" {

int fs_cholqr_work_size(int nmat, const fs_cholqr_desc *descs, int passes, int want_r,
                        int64_t *doubles) {
  return fs::guarded([&]() -> int {
    FS_CHECK_ARG(doubles, "cholqr: null output");
    Layout L;
    const int rc = make_layout(nmat, descs, passes, want_r != 0, L);
    if (rc != FS_OK) return rc;
    *doubles = L.total;
    return FS_OK;
  });
}

int fs_cholqr_f64(int nmat, const fs_cholqr_desc *descs, const double *Y, double *Q, double *Rt,
                  double *work, int64_t work_doubles, int passes, int shifted, int32_t *status,
                  void *stream) {
  return fs::guarded([&]() -> int {
    Layout L;
    const bool want_r = Rt != nullptr;
    int rc = make_layout(nmat, descs, passes, want_r, L);
    if (rc != FS_OK) return rc;
    if (nmat == 0) return FS_OK;
    FS_CHECK_ARG(Y && Q && work && status, "cholqr: null argument");
    FS_CHECK_ARG(work_doubles >= L.total, "cholqr: work buffer holds %lld doubles, %lld needed",
                 static_cast<long long>(work_doubles), static_cast<long long>(L.total));
    int64_t ngram = 0;
    for (int i = 0; i < nmat; ++i) ngram += L.nparts[i];
    std::vector<fs_gemm_group> gram(ngram), solve(nmat), acc(nmat);
    std::vector<fs_cholinv_desc> chol(nmat);
    std::vector<fs_triinv_desc> tri(nmat);
    for (int i = 0; i < nmat; ++i) {
      const int64_t l = descs[i].l;
      tri[i] = {l, L.r_base + L.l2_off[i], L.x_base + L.l2_off[i]};
    }
    // the current input of a pass: Y itself, then the packed output of the
    // pass before (row-major m x l in T0 / T1)
    const double *cur = Y;
    for (int p = 0; p < passes; ++p) {
      const bool last = p == passes - 1;
      const int64_t r_pass = L.r_base + (want_r ? p * L.l2_total : 0);
      double *out = last ? Q : work + L.t_base[p % 2];
      int64_t k = 0;
      for (int i = 0; i < nmat; ++i) {
        const fs_cholqr_desc &d = descs[i];
        const int64_t m = d.m, l = d.l;
        const int64_t c_off = p == 0 ? d.y_off : L.ml_off[i];
        const int64_t c_rs = p == 0 ? d.y_rs : l;
        const int64_t c_cs = p == 0 ? d.y_cs : 1;
        const int64_t np = L.nparts[i];
        for (int64_t part = 0; part < np; ++part, ++k) {
          const int64_t r0 = part * m / np, r1 = (part + 1) * m / np;
          const int64_t a_off = c_off + r0 * c_rs;
          gram[k] = {l, l, r1 - r0, a_off, c_cs, c_rs, a_off, c_rs, c_cs, L.g_off[i] + part * l * l, l};
        }
        const double ld = static_cast<double>(l);
        chol[i] = {l, L.g_base + L.g_off[i], l * l, np, r_pass + L.l2_off[i], 0,
                   shifted && p == 0 ? 11.0 * (static_cast<double>(m) * ld + ld * (ld + 1.0)) * DBL_EPSILON : 0.0};
        tri[i].r_off = r_pass + L.l2_off[i];
        solve[i] = {m, l, l, c_off, c_rs, c_cs, L.x_base + L.l2_off[i], l, 1,
                    last ? d.q_off : L.ml_off[i], l};
      }
      int32_t *st = status + static_cast<int64_t>(p) * nmat;
      const double *gin = p == 0 ? Y : cur;
      if ((rc = fs_grouped_gemm_f64(static_cast<int>(ngram), gram.data(), gin, gin, work, 1.0, 0.0, stream)) != FS_OK)
        return rc;
      if ((rc = fs_batched_cholinv_f64(nmat, chol.data(), work, work, nullptr, st, stream)) != FS_OK) return rc;
      if ((rc = fs_batched_triinv_f64(nmat, tri.data(), work, work, st, stream)) != FS_OK) return rc;
      if ((rc = fs_grouped_gemm_f64(nmat, solve.data(), cur, work, out, 1.0, 0.0, stream)) != FS_OK) return rc;
      if (want_r && p >= 1) {
        // running product R_p (R_{p-1} ... R_0), the last one into Rt
        const int64_t prev = p == 1 ? L.r_base : L.acc_base;
        for (int i = 0; i < nmat; ++i) {
          const int64_t l = descs[i].l;
          acc[i] = {l, l, l, r_pass + L.l2_off[i], l, 1, prev + L.l2_off[i], l, 1,
                    last ? descs[i].rt_off : L.acc_base + L.l2_off[i], l};
        }
        if ((rc = fs_grouped_gemm_f64(nmat, acc.data(), work, work, last ? Rt : work, 1.0, 0.0, stream)) != FS_OK)
          return rc;
      }
      cur = out;
    }
    return FS_OK;
  });
}

} // extern "C"
