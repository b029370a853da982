// Host-side planner: abelian charge fusion, block structure, permute index
// maps and the multiplet-aware truncation rule. Pure C++ (no HIP calls), so
// everything here can be exercised on a machine without a GPU.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <vector>

#include "fs_common.h"
#include "fs_plan.h"

namespace fs {

static thread_local std::string g_last_error;

void set_error(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

void clear_error() { g_last_error.clear(); }

// Kronecker sum of int8 charges, row-major (last leg fastest), int8
// wrap-around so that {0, -128} realises Z2.
// Follows u1_symmetric_tensor.py:8-19 (c1[j*n+k] = c0[j] + rs[k]).
std::vector<int8_t> fuse_charges(const std::vector<const int8_t *> &legs,
                                 const std::vector<int64_t> &sizes,
                                 const std::vector<uint8_t> &flip) {
  std::vector<int8_t> acc(1, 0);
  for (size_t l = 0; l < legs.size(); ++l) {
    const int64_t n = sizes[l];
    std::vector<int8_t> next(acc.size() * static_cast<size_t>(n));
    for (size_t j = 0; j < acc.size(); ++j) {
      for (int64_t k = 0; k < n; ++k) {
        int8_t c = legs[l][k];
        // int8 negation and addition both wrap (done in uint8 arithmetic)
        uint8_t cu = static_cast<uint8_t>(c);
        if (flip[l]) cu = static_cast<uint8_t>(0u - cu);
        next[j * n + k] = static_cast<int8_t>(
            static_cast<uint8_t>(static_cast<uint8_t>(acc[j]) + cu));
      }
    }
    acc.swap(next);
  }
  return acc;
}

// Kronecker sum of index contributions: out[j*n+k] = acc[j] + v_l[k].
static std::vector<int64_t>
kron_sum_i64(const std::vector<std::vector<int64_t>> &parts) {
  std::vector<int64_t> acc(1, 0);
  for (const auto &p : parts) {
    std::vector<int64_t> next(acc.size() * p.size());
    for (size_t j = 0; j < acc.size(); ++j)
      for (size_t k = 0; k < p.size(); ++k) next[j * p.size() + k] = acc[j] + p[k];
    acc.swap(next);
  }
  return acc;
}

// For every entry: its rank among entries of the same charge (stable order),
// plus per-charge counts. Equivalent to _numpy_get_indices
// (abelian_symmetric_tensor.py:78-100) whose mergesort keeps index order.
static void charge_ranks(const std::vector<int8_t> &ch, std::vector<int32_t> &rank,
                         std::array<int64_t, 256> &count) {
  count.fill(0);
  rank.resize(ch.size());
  for (size_t i = 0; i < ch.size(); ++i) {
    int c = static_cast<int>(ch[i]) + 128;
    rank[i] = static_cast<int32_t>(count[c]);
    count[c] += 1;
  }
}

fs_permute_plan *build_abelian_permute_plan(
    int ndim, const int64_t *shape, const int8_t *const *leg_charges,
    const uint8_t *signature, int old_nrr, int old_nblocks,
    const int8_t *old_block_irreps, const int64_t *old_block_offsets, int old_layout_f,
    int old_dense, const int32_t *axes, int new_nrr, int new_dense) {
  if (ndim < 2) throw std::invalid_argument("ndim must be >= 2");
  if (old_nrr < 1 || old_nrr >= ndim)
    throw std::invalid_argument("invalid old number of row legs");
  if (new_nrr < 1 || new_nrr >= ndim)
    throw std::invalid_argument("invalid new number of row legs");
  {
    std::vector<int> seen(ndim, 0);
    for (int i = 0; i < ndim; ++i) {
      if (axes[i] < 0 || axes[i] >= ndim || seen[axes[i]])
        throw std::invalid_argument("axes do not match tensor");
      seen[axes[i]] = 1;
    }
  }
  for (int l = 0; l < ndim; ++l)
    if (shape[l] < 1) throw std::invalid_argument("leg dimension must be >= 1");

  auto plan = new fs_permute_plan();
  try {
    // ---- old side ------------------------------------------------------
    std::vector<const int8_t *> rl(leg_charges, leg_charges + old_nrr);
    std::vector<const int8_t *> cl(leg_charges + old_nrr, leg_charges + ndim);
    std::vector<int64_t> rs(shape, shape + old_nrr), cs(shape + old_nrr, shape + ndim);
    std::vector<uint8_t> rf(signature, signature + old_nrr);
    std::vector<uint8_t> cf(ndim - old_nrr);
    for (int l = old_nrr; l < ndim; ++l) cf[l - old_nrr] = signature[l] ? 0 : 1;

    int64_t n_old_rows = 1, n_old_cols = 1;
    for (auto s : rs) n_old_rows *= s;
    for (auto s : cs) n_old_cols *= s;
    if (n_old_rows > std::numeric_limits<int32_t>::max() ||
        n_old_cols > std::numeric_limits<int32_t>::max())
      throw std::invalid_argument("matrix dimension exceeds int32 index range");

    plan->old_row_block.assign(n_old_rows, -1);
    plan->old_row_pos.assign(n_old_rows, 0);
    plan->old_col_block.assign(n_old_cols, -1);
    plan->old_col_pos.assign(n_old_cols, 0);

    if (old_dense) {
      for (int64_t r = 0; r < n_old_rows; ++r) {
        plan->old_row_block[r] = 0;
        plan->old_row_pos[r] = static_cast<int32_t>(r);
      }
      for (int64_t c = 0; c < n_old_cols; ++c) {
        plan->old_col_block[c] = 0;
        plan->old_col_pos[c] = static_cast<int32_t>(c);
      }
      plan->old_off = {0};
      plan->old_rs = {n_old_cols};
      plan->old_cs = {1};
      plan->old_ncoeff = n_old_rows * n_old_cols;
    } else {
      auto rch = fuse_charges(rl, rs, rf);
      auto cch = fuse_charges(cl, cs, cf);
      std::array<int64_t, 256> rcount, ccount;
      charge_ranks(rch, plan->old_row_pos, rcount);
      charge_ranks(cch, plan->old_col_pos, ccount);
      std::array<int32_t, 256> block_of;
      block_of.fill(-1);
      int prev = -1000;
      int64_t off = 0;
      for (int b = 0; b < old_nblocks; ++b) {
        int c = static_cast<int>(old_block_irreps[b]);
        if (c <= prev) throw std::invalid_argument("old block irreps must be sorted and unique");
        prev = c;
        int64_t nr = rcount[c + 128], nc = ccount[c + 128];
        if (nr == 0 || nc == 0)
          throw std::invalid_argument("old block irrep absent from row or column representation");
        block_of[c + 128] = b;
        plan->old_off.push_back(old_block_offsets ? old_block_offsets[b] : off);
        if (old_layout_f) {
          plan->old_rs.push_back(1);
          plan->old_cs.push_back(nr);
        } else {
          plan->old_rs.push_back(nc);
          plan->old_cs.push_back(1);
        }
        off += nr * nc;
      }
      plan->old_ncoeff = off;
      for (int64_t r = 0; r < n_old_rows; ++r)
        plan->old_row_block[r] = block_of[static_cast<int>(rch[r]) + 128];
      for (int64_t c = 0; c < n_old_cols; ++c)
        plan->old_col_block[c] = block_of[static_cast<int>(cch[c]) + 128];
    }

    // old dense strides of each leg: row legs within the row index, column
    // legs within the column index (row-major as in abelian:175-183)
    std::vector<int64_t> stride(ndim, 1);
    for (int l = old_nrr - 2; l >= 0; --l) stride[l] = stride[l + 1] * shape[l + 1];
    for (int l = ndim - 2; l >= old_nrr; --l) stride[l] = stride[l + 1] * shape[l + 1];

    // ---- new side ------------------------------------------------------
    std::vector<const int8_t *> nrl, ncl;
    std::vector<int64_t> nrs, ncs;
    std::vector<uint8_t> nrf, ncf;
    // contributions of each new leg to (old row index, old col index)
    std::vector<std::vector<int64_t>> row_partR, row_partC, col_partR, col_partC;
    for (int i = 0; i < ndim; ++i) {
      int L = axes[i];
      std::vector<int64_t> pR(shape[L], 0), pC(shape[L], 0);
      for (int64_t t = 0; t < shape[L]; ++t) {
        if (L < old_nrr) pR[t] = t * stride[L];
        else pC[t] = t * stride[L];
      }
      if (i < new_nrr) {
        nrl.push_back(leg_charges[L]);
        nrs.push_back(shape[L]);
        nrf.push_back(signature[L] ? 1 : 0);
        row_partR.push_back(std::move(pR));
        row_partC.push_back(std::move(pC));
      } else {
        ncl.push_back(leg_charges[L]);
        ncs.push_back(shape[L]);
        ncf.push_back(signature[L] ? 0 : 1);
        col_partR.push_back(std::move(pR));
        col_partC.push_back(std::move(pC));
      }
    }
    auto aR = kron_sum_i64(row_partR);
    auto aC = kron_sum_i64(row_partC);
    auto bR = kron_sum_i64(col_partR);
    auto bC = kron_sum_i64(col_partC);
    const int64_t n_new_rows = static_cast<int64_t>(aR.size());
    const int64_t n_new_cols = static_cast<int64_t>(bR.size());
    if (n_new_rows > std::numeric_limits<int32_t>::max() ||
        n_new_cols > std::numeric_limits<int32_t>::max())
      throw std::invalid_argument("matrix dimension exceeds int32 index range");

    std::vector<std::vector<int64_t>> rows_of, cols_of;  // per new block
    if (new_dense) {
      plan->new_irreps = {0};
      rows_of.resize(1);
      cols_of.resize(1);
      rows_of[0].resize(n_new_rows);
      cols_of[0].resize(n_new_cols);
      for (int64_t r = 0; r < n_new_rows; ++r) rows_of[0][r] = r;
      for (int64_t c = 0; c < n_new_cols; ++c) cols_of[0][c] = c;
    } else {
      auto rch = fuse_charges(nrl, nrs, nrf);
      auto cch = fuse_charges(ncl, ncs, ncf);
      std::array<int64_t, 256> rcount{}, ccount{};
      for (auto c : rch) rcount[static_cast<int>(c) + 128]++;
      for (auto c : cch) ccount[static_cast<int>(c) + 128]++;
      std::array<int32_t, 256> block_of;
      block_of.fill(-1);
      for (int c = 0; c < 256; ++c) {
        if (rcount[c] > 0 && ccount[c] > 0) {
          block_of[c] = static_cast<int32_t>(plan->new_irreps.size());
          plan->new_irreps.push_back(static_cast<int8_t>(c - 128));
        }
      }
      rows_of.resize(plan->new_irreps.size());
      cols_of.resize(plan->new_irreps.size());
      for (int64_t r = 0; r < n_new_rows; ++r) {
        int b = block_of[static_cast<int>(rch[r]) + 128];
        if (b >= 0) rows_of[b].push_back(r);
      }
      for (int64_t c = 0; c < n_new_cols; ++c) {
        int b = block_of[static_cast<int>(cch[c]) + 128];
        if (b >= 0) cols_of[b].push_back(c);
      }
    }

    const size_t nb = plan->new_irreps.size();
    int64_t off = 0, rstart = 0, cstart = 0;
    for (size_t b = 0; b < nb; ++b) {
      const int64_t nr = static_cast<int64_t>(rows_of[b].size());
      const int64_t nc = static_cast<int64_t>(cols_of[b].size());
      plan->new_shapes.push_back(nr);
      plan->new_shapes.push_back(nc);
      plan->new_off.push_back(off);
      plan->new_rowstart.push_back(rstart);
      plan->new_colstart.push_back(cstart);
      for (auto r : rows_of[b]) {
        plan->row_a.push_back(static_cast<int32_t>(aR[r]));
        plan->row_a.push_back(static_cast<int32_t>(aC[r]));
      }
      for (auto c : cols_of[b]) {
        plan->col_b.push_back(static_cast<int32_t>(bR[c]));
        plan->col_b.push_back(static_cast<int32_t>(bC[c]));
      }
      off += nr * nc;
      rstart += nr;
      cstart += nc;
    }
    plan->new_ncoeff = off;
    plan->n_old_rows = n_old_rows;
    plan->n_old_cols = n_old_cols;
    plan->n_new_rows_total = rstart;
    plan->n_new_cols_total = cstart;
    plan->finalize_tiles();
  } catch (...) {
    delete plan;
    throw;
  }
  return plan;
}

} // namespace fs

void fs_permute_plan::finalize_tiles() {
  tile_prefix.assign(1, 0);
  const size_t nb = new_irreps.size();
  for (size_t b = 0; b < nb; ++b) {
    int64_t n = new_shapes[2 * b] * new_shapes[2 * b + 1];
    tile_prefix.push_back(tile_prefix.back() + (n + kPermuteTile - 1) / kPermuteTile);
  }
}


namespace {

// Row-major outer product of the degeneracy vectors of `legs` (last leg
// fastest): the degeneracy of every irreps combination of those legs.
std::vector<int64_t> degen_outer(const std::vector<int64_t> &legs, const int64_t *degen,
                                 const std::vector<int64_t> &doff, const int64_t *nirr) {
  std::vector<int64_t> out(1, 1);
  for (int64_t leg : legs) {
    const int64_t n = nirr[leg];
    std::vector<int64_t> next(out.size() * static_cast<size_t>(n));
    for (size_t i = 0; i < out.size(); ++i)
      for (int64_t j = 0; j < n; ++j) next[i * n + j] = out[i] * degen[doff[leg] + j];
    out.swap(next);
  }
  return out;
}

// start[c][t]: first row (column) of combination c inside target t, i.e. the
// exclusive running sum over c of ed[c] * id[c][t]; total[t] is the full sum.
void block_starts(const std::vector<int64_t> &ed, const int64_t *id, int64_t nt,
                  std::vector<int64_t> &start, std::vector<int64_t> &total) {
  start.assign(ed.size() * static_cast<size_t>(nt), 0);
  total.assign(static_cast<size_t>(nt), 0);
  for (int64_t t = 0; t < nt; ++t) {
    int64_t run = 0;
    for (size_t c = 0; c < ed.size(); ++c) {
      start[c * nt + t] = run;
      run += ed[c] * id[c * nt + t];
    }
    total[t] = run;
  }
}

} // namespace

extern "C" {

const char *fs_last_error(void) { return fs::g_last_error.c_str(); }

int fs_abi_version(void) { return 1; }

int fs_u1_fuse(int nlegs, const int8_t *const *legs, const int64_t *sizes,
               const uint8_t *flip, int8_t *out) {
  return fs::guarded([&]() -> int {
    FS_CHECK_ARG(nlegs >= 1 && legs && sizes && flip && out, "fs_u1_fuse: invalid arguments");
    std::vector<const int8_t *> l(legs, legs + nlegs);
    std::vector<int64_t> s(sizes, sizes + nlegs);
    std::vector<uint8_t> f(flip, flip + nlegs);
    auto r = fs::fuse_charges(l, s, f);
    std::memcpy(out, r.data(), r.size());
    return FS_OK;
  });
}

// Greedy global truncation, same decision sequence as
// _numba_find_block_cuts (misc_tools/svd_tools.py:57-74): repeatedly take the
// block holding the current largest value (first index on ties) and keep
// every value >= degen_ratio * that value before re-checking the stop rule.
int fs_find_block_cuts(int nblocks, const double *const *values,
                       const int64_t *sizes, int64_t cut_target,
                       const int64_t *dims, double cutoff, double degen_ratio,
                       int64_t *cuts) {
  return fs::guarded([&]() -> int {
    FS_CHECK_ARG(nblocks >= 0 && cuts, "fs_find_block_cuts: invalid arguments");
    FS_CHECK_ARG(degen_ratio <= 1.0, "fs_find_block_cuts: degen_ratio must be <= 1");
    if (nblocks == 0) return FS_OK;
    std::vector<double> maxv(nblocks);
    for (int b = 0; b < nblocks; ++b) {
      cuts[b] = 0;
      maxv[b] = sizes[b] > 0 ? values[b][0] : -1.0;
    }
    auto argmax = [&]() {
      int bi = 0;
      for (int b = 1; b < nblocks; ++b)
        if (maxv[b] > maxv[bi]) bi = b;
      return bi;
    };
    int64_t kept = 0;
    int bi = argmax();
    while (maxv[bi] > cutoff && kept < cut_target) {
      const double c = degen_ratio * maxv[bi];
      while (maxv[bi] >= c) {
        cuts[bi] += 1;
        kept += dims ? dims[bi] : 1;
        if (cuts[bi] < sizes[bi]) maxv[bi] = values[bi][cuts[bi]];
        else maxv[bi] = -1.0;
        bi = argmax();
      }
    }
    return FS_OK;
  });
}

int fs_abelian_permute_plan_create(int ndim, const int64_t *shape,
                                   const int8_t *const *leg_charges,
                                   const uint8_t *signature, int old_nrr,
                                   int old_nblocks,
                                   const int8_t *old_block_irreps,
                                   const int64_t *old_block_offsets,
                                   int old_layout_f, int old_dense,
                                   const int32_t *axes, int new_nrr,
                                   int new_dense, fs_permute_plan **plan) {
  return fs::guarded([&]() -> int {
    FS_CHECK_ARG(shape && leg_charges && signature && axes && plan,
                 "fs_abelian_permute_plan_create: null argument");
    FS_CHECK_ARG(old_dense || old_nblocks == 0 || old_block_irreps,
                 "fs_abelian_permute_plan_create: missing old block irreps");
    *plan = fs::build_abelian_permute_plan(
        ndim, shape, leg_charges, signature, old_nrr, old_nblocks,
        old_block_irreps, old_block_offsets, old_layout_f, old_dense, axes, new_nrr, new_dense);
    return FS_OK;
  });
}

int fs_permute_plan_info(const fs_permute_plan *plan, int64_t *new_nblocks,
                         int64_t *new_ncoeff, int64_t *old_ncoeff) {
  return fs::guarded([&]() -> int {
    FS_CHECK_ARG(plan, "fs_permute_plan_info: null plan");
    if (new_nblocks) *new_nblocks = static_cast<int64_t>(plan->new_irreps.size());
    if (new_ncoeff) *new_ncoeff = plan->new_ncoeff;
    if (old_ncoeff) *old_ncoeff = plan->old_ncoeff;
    return FS_OK;
  });
}

int fs_permute_plan_blocks(const fs_permute_plan *plan, int8_t *irreps,
                           int64_t *shapes) {
  return fs::guarded([&]() -> int {
    FS_CHECK_ARG(plan, "fs_permute_plan_blocks: null plan");
    if (irreps) std::memcpy(irreps, plan->new_irreps.data(), plan->new_irreps.size());
    if (shapes)
      std::memcpy(shapes, plan->new_shapes.data(), plan->new_shapes.size() * sizeof(int64_t));
    return FS_OK;
  });
}

int fs_permute_plan_tables(const fs_permute_plan *plan, int64_t *sizes,
                           int32_t *old_row_block, int32_t *old_row_pos,
                           int32_t *old_col_block, int32_t *old_col_pos,
                           int32_t *row_a, int32_t *col_b, int64_t *old_off,
                           int64_t *old_rs, int64_t *old_cs) {
  return fs::guarded([&]() -> int {
    FS_CHECK_ARG(plan, "fs_permute_plan_tables: null plan");
    if (sizes) {
      sizes[0] = plan->n_old_rows;
      sizes[1] = plan->n_old_cols;
      sizes[2] = plan->n_new_rows_total;
      sizes[3] = plan->n_new_cols_total;
      sizes[4] = static_cast<int64_t>(plan->old_off.size());
    }
    auto cp = [](int32_t *dst, const std::vector<int32_t> &v) {
      if (dst) std::memcpy(dst, v.data(), v.size() * sizeof(int32_t));
    };
    auto cp64 = [](int64_t *dst, const std::vector<int64_t> &v) {
      if (dst) std::memcpy(dst, v.data(), v.size() * sizeof(int64_t));
    };
    cp(old_row_block, plan->old_row_block);
    cp(old_row_pos, plan->old_row_pos);
    cp(old_col_block, plan->old_col_block);
    cp(old_col_pos, plan->old_col_pos);
    cp(row_a, plan->row_a);
    cp(col_b, plan->col_b);
    cp64(old_off, plan->old_off);
    cp64(old_rs, plan->old_rs);
    cp64(old_cs, plan->old_cs);
    return FS_OK;
  });
}

int fs_lie_plan_tables(const fs_lie_pattern *pat, int64_t ndegen, const int64_t *degen,
                       const int64_t *in_block, int64_t nblocks, const int64_t *block_off,
                       const int64_t *block_rs, const int64_t *block_cs, int64_t *tasks,
                       int64_t nrows, int64_t *rows, int64_t *digits, int64_t *shapes) {
  return fs::guarded([&]() -> int {
    FS_CHECK_ARG(pat && degen && in_block && tasks && rows && digits && shapes,
                 "lie plan tables: null argument");
    const int64_t nd = pat->ndim, nrr = pat->nrr, nro = pat->nrr_out, ns = pat->ntasks;
    const int64_t nti = pat->nt_in, nto = pat->nt_out;
    FS_CHECK_ARG(nd >= 1 && nd <= FS_LIE_MAX_LEGS && nrr >= 0 && nrr <= nd && nro >= 0 && nro <= nd,
                 "lie plan tables: bad leg counts");
    FS_CHECK_ARG(ns >= 0 && nti >= 0 && nto >= 0 && nblocks >= 0, "lie plan tables: negative size");
    FS_CHECK_ARG(nblocks == 0 || (block_off && block_rs && block_cs), "lie plan tables: null block table");
    std::vector<int64_t> doff(static_cast<size_t>(nd) + 1, 0);
    std::vector<bool> seen(static_cast<size_t>(nd), false);
    for (int64_t k = 0; k < nd; ++k) {
      FS_CHECK_ARG(pat->nirr[k] >= 1, "lie plan tables: leg %lld has no irrep", static_cast<long long>(k));
      doff[k + 1] = doff[k] + pat->nirr[k];
      const int64_t a = pat->axes[k];
      FS_CHECK_ARG(a >= 0 && a < nd && !seen[a], "lie plan tables: axes is not a permutation");
      seen[a] = true;
    }
    FS_CHECK_ARG(doff[nd] == ndegen, "lie plan tables: %lld degeneracies for %lld irreps",
                 static_cast<long long>(ndegen), static_cast<long long>(doff[nd]));
    std::vector<int64_t> legs_ri, legs_ci, legs_ro, legs_co;
    for (int64_t k = 0; k < nd; ++k) {
      (k < nrr ? legs_ri : legs_ci).push_back(k);
      (k < nro ? legs_ro : legs_co).push_back(pat->axes[k]);
    }
    const std::vector<int64_t> edr = degen_outer(legs_ri, degen, doff, pat->nirr);
    const std::vector<int64_t> edc = degen_outer(legs_ci, degen, doff, pat->nirr);
    const std::vector<int64_t> edr2 = degen_outer(legs_ro, degen, doff, pat->nirr);
    const std::vector<int64_t> edc2 = degen_outer(legs_co, degen, doff, pat->nirr);
    std::vector<int64_t> rstart, cstart, rstart2, cstart2, rtot, ctot, nrows_t, ncols_t;
    block_starts(edr, pat->idr_in, nti, rstart, rtot);
    block_starts(edc, pat->idc_in, nti, cstart, ctot);
    block_starts(edr2, pat->idr_out, nto, rstart2, nrows_t);
    block_starts(edc2, pat->idc_out, nto, cstart2, ncols_t);
    std::vector<int64_t> noff(static_cast<size_t>(nto), 0);
    for (int64_t t = 0, run = 0; t < nto; ++t) {
      shapes[2 * t] = nrows_t[t];
      shapes[2 * t + 1] = ncols_t[t];
      noff[t] = run;
      run += nrows_t[t] * ncols_t[t];
    }
    for (int64_t t = 0; t < nti; ++t)
      FS_CHECK_ARG(in_block[t] >= -1 && in_block[t] < nblocks, "lie plan tables: bad stored block");

    int64_t row0 = 0, ein = 0, eout = 0;
    int64_t D[FS_LIE_MAX_LEGS], Dn[FS_LIE_MAX_LEGS], so[FS_LIE_MAX_LEGS], sn[FS_LIE_MAX_LEGS];
    for (int64_t s = 0; s < ns; ++s) {
      int64_t E = 1;
      for (int64_t k = 0; k < nd; ++k) {
        const int64_t i = pat->sig_idx[s * nd + k];
        FS_CHECK_ARG(i >= 0 && i < pat->nirr[k], "lie plan tables: irrep index out of range");
        D[k] = degen[doff[k] + i];
        E *= D[k];
      }
      const int64_t ir = pat->ir_in[s], ic = pat->ic_in[s];
      const int64_t ir2 = pat->ir_out[s], ic2 = pat->ic_out[s];
      FS_CHECK_ARG(ir >= 0 && ir < static_cast<int64_t>(edr.size()) && ic >= 0 &&
                       ic < static_cast<int64_t>(edc.size()) && ir2 >= 0 &&
                       ir2 < static_cast<int64_t>(edr2.size()) && ic2 >= 0 &&
                       ic2 < static_cast<int64_t>(edc2.size()),
                   "lie plan tables: irreps combination out of range");
      const int64_t nin = pat->n_in[s], nout = pat->n_out[s];
      FS_CHECK_ARG(nin >= 0 && nout >= 0 && row0 + nin + nout <= nrows,
                   "lie plan tables: row table too short");
      for (int64_t j = 0; j < nin; ++j, ++ein) {
        const int64_t t = pat->in_t[ein];
        FS_CHECK_ARG(t >= 0 && t < nti, "lie plan tables: input target out of range");
        const int64_t row = rstart[ir * nti + t] + pat->in_a[ein] * edr[ir];
        const int64_t col = cstart[ic * nti + t] + pat->in_b[ein] * edc[ic];
        const int64_t b = in_block[t];
        int64_t *r = rows + 3 * (row0 + j);
        if (b >= 0) {
          r[0] = block_off[b] + row * block_rs[b] + col * block_cs[b];
          r[1] = block_rs[b];
          r[2] = block_cs[b];
        } else {
          r[0] = -1;
          r[1] = r[2] = 0;
        }
      }
      for (int64_t j = 0; j < nout; ++j, ++eout) {
        const int64_t t = pat->out_t[eout];
        FS_CHECK_ARG(t >= 0 && t < nto, "lie plan tables: output target out of range");
        int64_t *r = rows + 3 * (row0 + nin + j);
        r[0] = noff[t] + (rstart2[ir2 * nto + t] + pat->out_a[eout] * edr2[ir2]) * ncols_t[t] +
               cstart2[ic2 * nto + t] + pat->out_b[eout] * edc2[ic2];
        r[1] = ncols_t[t];
        r[2] = 1;
      }
      // digits: one per NEW leg (the external index runs row-major in the new
      // order), with its strides on the old and the new side
      for (int64_t k = nd - 1, run_r = 1, run_c = 1; k >= 0; --k) {  // old strides
        if (k >= nrr) { so[k] = run_c; run_c *= D[k]; }
        else { so[k] = run_r; run_r *= D[k]; }
      }
      for (int64_t k = 0; k < nd; ++k) Dn[k] = D[pat->axes[k]];
      for (int64_t k = nd - 1, run_r = 1, run_c = 1; k >= 0; --k) {  // new strides
        if (k >= nro) { sn[k] = run_c; run_c *= Dn[k]; }
        else { sn[k] = run_r; run_r *= Dn[k]; }
      }
      for (int64_t k = 0; k < nd; ++k) {
        const int64_t a = pat->axes[k];
        int64_t *d = digits + 5 * (s * nd + k);
        d[0] = Dn[k];
        d[1] = a < nrr ? so[a] : 0;
        d[2] = a >= nrr ? so[a] : 0;
        d[3] = k < nro ? sn[k] : 0;
        d[4] = k >= nro ? sn[k] : 0;
      }
      int64_t *tk = tasks + 8 * s;
      tk[0] = E;
      tk[1] = pat->w_off[s];
      tk[2] = nin;
      tk[3] = nout;
      tk[4] = pat->n_out_pad[s];
      tk[5] = row0;
      tk[6] = nd;
      tk[7] = s * nd;
      row0 += nin + nout;
    }
    return FS_OK;
  });
}

} // extern "C"
