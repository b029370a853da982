// Host-side planner: abelian charge fusion, block structure, permute index
// maps and the multiplet-aware truncation rule. Pure C++ (no HIP calls), so
// everything here can be exercised on a machine without a GPU.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
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

// ---- abelian permute as boxes ------------------------------------------------
//
// The reference permutes coefficient by coefficient (_numba_abelian_transpose,
// abelian_symmetric_tensor.py:103-250): old (block, row, col) -> dense index ->
// new (block, row, col), where the rows of a block are the dense row indices of
// its charge in increasing order (_numpy_get_indices, :78-100). Here the same
// map is cut into boxes: a box fixes one run of equal charges on every leg.
// For a row multi-index (i_0 .. i_{k-1}) of charge Q, its rank among the rows
// of charge Q is
//     sum_p sum_{t < i_p} S_{p+1}(Q - P_p - q_p(t)),
// S_p(x) the number of index tuples of legs p..k-1 with total charge x and P_p
// the charge of the prefix. Within a box every q_p(t) for t in the run is
// constant, so the rank is base + sum_p (i_p - start_p) * S_{p+1}(Q - P_p - q_p):
// affine in the digits, on the old and on the new side alike.

namespace {

struct Run {
  int64_t start, len;
  uint8_t q;  // raw int8 charge bits
};

std::vector<Run> leg_runs(const int8_t *c, int64_t n) {
  std::vector<Run> out;
  for (int64_t t = 0; t < n; ++t) {
    const uint8_t q = static_cast<uint8_t>(c[t]);
    if (!out.empty() && out.back().q == q) ++out.back().len;
    else out.push_back({t, 1, q});
  }
  return out;
}

inline uint8_t signed_charge(uint8_t q, bool flip) {
  return flip ? static_cast<uint8_t>(0u - q) : q;  // int8 negation wraps
}

// All run tuples of one side of a matrix view (legs in order, row-major over
// the legs' run indices, last leg fastest).
struct Side {
  std::vector<int> legs;
  std::vector<int64_t> nruns;
  std::array<int64_t, 256> count{};  // index tuples per fused charge
  int64_t ntuples = 1;
  std::vector<uint8_t> q;       // per tuple: fused signed charge
  std::vector<int64_t> base;    // per tuple: in-block position (dense: dense rank) of its first element
  std::vector<int64_t> stride;  // per tuple and leg: position step per digit
  std::vector<int64_t> runidx;  // per tuple and leg: run index
};

Side make_side(const std::vector<int> &legs, const std::vector<uint8_t> &flip,
               const std::vector<std::vector<Run>> &runs, const int64_t *shape, bool dense) {
  Side s;
  s.legs = legs;
  const size_t k = legs.size();
  for (int l : legs) s.nruns.push_back(static_cast<int64_t>(runs[l].size()));
  for (int64_t n : s.nruns) s.ntuples *= n;
  // suffix counts S_p, p = k .. 0
  std::vector<std::array<int64_t, 256>> S(k + 1);
  S[k].fill(0);
  S[k][0] = 1;
  for (size_t p = k; p-- > 0;) {
    std::array<int64_t, 256> M{};
    for (const Run &r : runs[legs[p]]) M[signed_charge(r.q, flip[p])] += r.len;
    S[p].fill(0);
    for (int a = 0; a < 256; ++a)
      if (M[a])
        for (int x = 0; x < 256; ++x)
          if (S[p + 1][x]) S[p][(a + x) & 255] += M[a] * S[p + 1][x];
  }
  s.count = S[0];
  // per leg and run: signed charge counts of the states before the run
  std::vector<std::vector<std::vector<std::pair<uint8_t, int64_t>>>> before(k);
  for (size_t p = 0; p < k; ++p) {
    std::array<int64_t, 256> seen{};
    for (const Run &r : runs[legs[p]]) {
      std::vector<std::pair<uint8_t, int64_t>> snap;
      for (int a = 0; a < 256; ++a)
        if (seen[a]) snap.push_back({static_cast<uint8_t>(a), seen[a]});
      before[p].push_back(std::move(snap));
      seen[signed_charge(r.q, flip[p])] += r.len;
    }
  }
  std::vector<int64_t> dstride(k, 1);
  for (size_t p = k; p-- > 1;) dstride[p - 1] = dstride[p] * shape[legs[p]];
  s.q.resize(s.ntuples);
  s.base.resize(s.ntuples);
  s.stride.resize(s.ntuples * k);
  s.runidx.resize(s.ntuples * k);
  std::vector<int64_t> ri(k, 0);
  for (int64_t t = 0; t < s.ntuples; ++t) {
    int64_t rem = t;
    for (size_t p = k; p-- > 0;) {
      ri[p] = rem % s.nruns[p];
      rem /= s.nruns[p];
    }
    unsigned Q = 0;
    for (size_t p = 0; p < k; ++p) Q += signed_charge(runs[legs[p]][ri[p]].q, flip[p]);
    s.q[t] = static_cast<uint8_t>(Q);
    int64_t base = 0;
    unsigned P = 0;
    for (size_t p = 0; p < k; ++p) {
      const Run &r = runs[legs[p]][ri[p]];
      s.runidx[t * k + p] = ri[p];
      if (dense) {
        base += r.start * dstride[p];
        s.stride[t * k + p] = dstride[p];
        continue;
      }
      const uint8_t qp = signed_charge(r.q, flip[p]);
      for (const auto &[qq, cnt] : before[p][ri[p]]) base += cnt * S[p + 1][(Q - P - qq) & 255];
      s.stride[t * k + p] = S[p + 1][(Q - P - qp) & 255];
      P += qp;
    }
    s.base[t] = base;
  }
  return s;
}

struct Digit {
  int64_t len, s, d, l;  // length, source stride, destination stride, staging stride
};

uint32_t magic_of(int64_t len) {
  return static_cast<uint32_t>(((uint64_t(1) << 32) + static_cast<uint64_t>(len) - 1) /
                               static_cast<uint64_t>(len));
}

// Record of one box: digits given in destination (new leg) order, slowest
// first. The slowest digits are kept out of the tiles (one index per tile)
// until the remaining volume fits kPermuteTileVol; the next digit is cut in
// chunks of kPermuteTileVol / (volume of the faster digits) when that is at
// least 2 (else it is a slow digit too). The tile's destination-order list
// (fastest first, contiguous neighbours merged) ends with the cut digit; the
// source-order list sorts the digits by source stride and merges neighbours
// contiguous in both the source and the staging buffer, never onto an entry
// holding a cut digit (its length differs in the last chunk). Checks the whole
// box against both buffers and the 32-bit in-tile offsets.
constexpr int kK = kPermuteMaxDigits;

void make_box(const Digit *dig0, int ndig0, int64_t src_base, int64_t dst_base, int64_t src_limit,
              int64_t dst_limit, PermBox &b) {
  Digit dig[kK];
  int n = 0;
  for (int i = 0; i < ndig0; ++i)
    if (dig0[i].len > 1) {
      if (n == kK) throw std::invalid_argument("permute plan: too many digits in a box");
      dig[n++] = dig0[i];
    }
  int64_t vol = 1, slast = 0, dlast = 0;
  for (int i = 0; i < n; ++i) {
    vol *= dig[i].len;
    slast += (dig[i].len - 1) * dig[i].s;
    dlast += (dig[i].len - 1) * dig[i].d;
  }
  if (src_base >= 0 && src_base + slast >= src_limit)
    throw std::invalid_argument("permute plan: box reads outside the source");
  if (dst_base < 0 || dst_base + dlast >= dst_limit)
    throw std::invalid_argument("permute plan: box writes outside the destination");
  std::memset(&b, 0, sizeof(b));
  b.src_base = src_base >= 0 ? src_base : -1;
  b.dst_base = dst_base;
  b.s_cut = -1;
  // slow digits and cut digit
  int nslow = 0, cut = -1;
  int64_t chunk = 0;
  if (vol > kPermuteTileVol) {
    int64_t rest = vol;
    for (int i = 0; i < n; ++i) {
      rest /= dig[i].len;
      if (rest <= kPermuteTileVol) {
        const int64_t c = kPermuteTileVol / rest;  // < dig[i].len
        if (c >= 2) {
          cut = i;
          chunk = c;
          nslow = i;
        } else {
          nslow = i + 1;
        }
        break;
      }
    }
  }
  b.nslow = nslow;
  for (int k = 0; k < nslow; ++k) {  // fastest slow digit first
    const Digit &g = dig[nslow - 1 - k];
    b.slow_len[k] = static_cast<int32_t>(g.len);
    b.slow_s[k] = g.s;
    b.slow_d[k] = g.d;
  }
  // tile digits (destination order, slowest first), the cut one at its chunk length
  Digit td[kK];
  int nt = 0;
  for (int i = nslow; i < n; ++i) {
    td[nt] = dig[i];
    if (i == cut) td[nt].len = chunk;
    ++nt;
  }
  if (cut >= 0) {
    b.nchunks = static_cast<int32_t>((dig[cut].len + chunk - 1) / chunk);
    b.chunk = static_cast<int32_t>(chunk);
    b.cut_s = dig[cut].s;
    b.cut_d = dig[cut].d;
  } else {
    b.nchunks = 1;
    b.chunk = 0;
  }
  const int64_t edge = cut >= 0 ? dig[cut].len - (b.nchunks - 1) * chunk : 0;  // last chunk's length
  int64_t tvol = 1, tslast = 0, tdlast = 0;
  for (int i = nt; i-- > 0;) {
    td[i].l = tvol;  // staging: destination order, row-major
    tvol *= td[i].len;
    tslast += (td[i].len - 1) * td[i].s;
    tdlast += (td[i].len - 1) * td[i].d;
  }
  if (tslast >= (int64_t(1) << 31) || tdlast >= (int64_t(1) << 31))
    throw std::invalid_argument("permute plan: block too large for 32-bit in-tile offsets");
  b.vol = static_cast<int32_t>(tvol);
  b.vol_edge = cut >= 0 ? static_cast<int32_t>(tvol / chunk * edge) : b.vol;
  // destination list, fastest first; the cut digit (td[0] when cut) is the last entry
  int nd = 0;
  int64_t dl[kK], ds[kK];
  for (int i = nt; i-- > 0;) {
    if (nd > 0 && td[i].d == dl[nd - 1] * ds[nd - 1]) {
      dl[nd - 1] *= td[i].len;
    } else {
      dl[nd] = td[i].len;
      ds[nd] = td[i].d;
      ++nd;
    }
  }
  // source list: by increasing source stride (stable), merged where contiguous
  int order[kK];
  for (int i = 0; i < nt; ++i) order[i] = i;
  for (int i = 1; i < nt; ++i)  // insertion sort: stable, at most kK entries
    for (int j = i; j > 0 && td[order[j]].s < td[order[j - 1]].s; --j) std::swap(order[j], order[j - 1]);
  int ns = 0;
  int64_t sl[kK], ss[kK], sls[kK], sfac[kK];
  bool scut[kK];
  for (int q = 0; q < nt; ++q) {
    const Digit &g = td[order[q]];
    const bool is_cut = cut >= 0 && order[q] == 0;
    if (ns > 0 && !scut[ns - 1] && g.s == sl[ns - 1] * ss[ns - 1] && g.l == sl[ns - 1] * sls[ns - 1]) {
      sl[ns - 1] *= g.len;
      if (is_cut) scut[ns - 1] = true;
      else sfac[ns - 1] *= g.len;
    } else {
      sl[ns] = g.len;
      ss[ns] = g.s;
      sls[ns] = g.l;
      scut[ns] = is_cut;
      sfac[ns] = is_cut ? 1 : g.len;
      ++ns;
    }
  }
  b.ns = ns;
  b.nd = nd;
  for (int k = 0; k < ns; ++k) {
    b.slen[k] = static_cast<int32_t>(sl[k]);
    b.smag[k] = magic_of(sl[k]);
    b.sstr[k] = static_cast<int32_t>(ss[k]);
    b.lstr[k] = static_cast<int32_t>(sls[k]);
    // the last chunk changes the length of the entry holding the cut digit
    if (scut[k] && edge != chunk) {
      b.s_cut = k;
      b.s_len_edge = static_cast<int32_t>(sfac[k] * edge);
      b.s_drop_edge = b.s_len_edge == 1;
      b.s_mag_edge = b.s_len_edge > 1 ? magic_of(b.s_len_edge) : 0u;
    }
  }
  for (int k = 0; k < nd; ++k) {
    b.dlen[k] = static_cast<int32_t>(dl[k]);
    b.dmag[k] = magic_of(dl[k]);
    b.dstr[k] = static_cast<int32_t>(ds[k]);
  }
  // the cut digit is in the last destination entry (its length is not used
  // to decode, the tile volume bounds the walk; kept exact for the record)
  b.d_len_edge = cut >= 0 ? static_cast<int32_t>(dl[nd - 1] / chunk * edge) : b.dlen[nd > 0 ? nd - 1 : 0];
}
}  // namespace

fs_permute_plan *build_abelian_permute_plan(
    int ndim, const int64_t *shape, const int8_t *const *leg_charges,
    const uint8_t *signature, int old_nrr, int old_nblocks,
    const int8_t *old_block_irreps, const int64_t *old_block_offsets, int old_layout_f,
    int old_dense, const int32_t *axes, int new_nrr, int new_dense) {
  if (ndim < 2) throw std::invalid_argument("ndim must be >= 2");
  if (old_nrr < 1 || old_nrr >= ndim) throw std::invalid_argument("invalid old number of row legs");
  if (new_nrr < 1 || new_nrr >= ndim) throw std::invalid_argument("invalid new number of row legs");
  {
    std::vector<int> seen(ndim, 0);
    for (int i = 0; i < ndim; ++i) {
      if (axes[i] < 0 || axes[i] >= ndim || seen[axes[i]]) throw std::invalid_argument("axes do not match tensor");
      seen[axes[i]] = 1;
    }
  }
  for (int l = 0; l < ndim; ++l)
    if (shape[l] < 1) throw std::invalid_argument("leg dimension must be >= 1");

  auto plan = std::make_unique<fs_permute_plan>();
  std::vector<std::vector<Run>> runs(ndim);
  for (int l = 0; l < ndim; ++l) runs[l] = leg_runs(leg_charges[l], shape[l]);

  // sides: rows fuse with the signature, columns with its opposite
  std::vector<int> orl, ocl, nrl, ncl;
  std::vector<uint8_t> orf, ocf, nrf, ncf;
  for (int l = 0; l < ndim; ++l) {
    if (l < old_nrr) {
      orl.push_back(l);
      orf.push_back(signature[l] ? 1 : 0);
    } else {
      ocl.push_back(l);
      ocf.push_back(signature[l] ? 0 : 1);
    }
  }
  for (int i = 0; i < ndim; ++i) {
    const int L = axes[i];
    if (i < new_nrr) {
      nrl.push_back(L);
      nrf.push_back(signature[L] ? 1 : 0);
    } else {
      ncl.push_back(L);
      ncf.push_back(signature[L] ? 0 : 1);
    }
  }
  const Side orow = make_side(orl, orf, runs, shape, old_dense != 0);
  const Side ocol = make_side(ocl, ocf, runs, shape, old_dense != 0);
  const Side nrow = make_side(nrl, nrf, runs, shape, new_dense != 0);
  const Side ncol = make_side(ncl, ncf, runs, shape, new_dense != 0);

  // old blocks: offset and (row, column) strides per charge
  std::array<int64_t, 256> o_off{}, o_rs{}, o_cs{};
  std::array<bool, 256> o_present{};
  int64_t src_limit = 0;
  if (old_dense) {
    int64_t nr = 1, nc = 1;
    for (int l : orl) nr *= shape[l];
    for (int l : ocl) nc *= shape[l];
    o_present.fill(true);
    o_off.fill(0);
    o_rs.fill(nc);
    o_cs.fill(1);
    src_limit = nr * nc;
    plan->old_ncoeff = nr * nc;
  } else {
    int prev = -1000;
    int64_t off = 0;
    for (int b = 0; b < old_nblocks; ++b) {
      const int c = static_cast<int>(old_block_irreps[b]);
      if (c <= prev) throw std::invalid_argument("old block irreps must be sorted and unique");
      prev = c;
      const uint8_t u = static_cast<uint8_t>(old_block_irreps[b]);
      const int64_t nr = orow.count[u], nc = ocol.count[u];
      if (nr == 0 || nc == 0) throw std::invalid_argument("old block irrep absent from row or column representation");
      o_present[u] = true;
      o_off[u] = old_block_offsets ? old_block_offsets[b] : off;
      if (o_off[u] < 0) throw std::invalid_argument("negative old block offset");
      o_rs[u] = old_layout_f ? 1 : nc;
      o_cs[u] = old_layout_f ? nr : 1;
      src_limit = std::max(src_limit, o_off[u] + nr * nc);
      off += nr * nc;
    }
    plan->old_ncoeff = old_block_offsets ? src_limit : off;
  }

  // new blocks: charges present on both new sides, ascending as int8
  std::array<int64_t, 256> n_off{}, n_cols{};
  std::array<bool, 256> n_present{};
  int64_t noff = 0;
  if (new_dense) {
    int64_t nr = 1, nc = 1;
    for (int l : nrl) nr *= shape[l];
    for (int l : ncl) nc *= shape[l];
    plan->new_irreps = {0};
    plan->new_shapes = {nr, nc};
    n_present.fill(true);
    n_cols.fill(nc);
    noff = nr * nc;
  } else {
    for (int c = -128; c < 128; ++c) {
      const uint8_t u = static_cast<uint8_t>(c);
      if (nrow.count[u] > 0 && ncol.count[u] > 0) {
        plan->new_irreps.push_back(static_cast<int8_t>(c));
        plan->new_shapes.push_back(nrow.count[u]);
        plan->new_shapes.push_back(ncol.count[u]);
        n_present[u] = true;
        n_off[u] = noff;
        n_cols[u] = ncol.count[u];
        noff += nrow.count[u] * ncol.count[u];
      }
    }
  }
  plan->new_ncoeff = noff;

  // old column tuples by charge
  std::vector<std::vector<int64_t>> by_q(256);
  for (int64_t b = 0; b < ocol.ntuples; ++b) by_q[ocol.q[b]].push_back(b);
  std::vector<int64_t> all_cols(static_cast<size_t>(ocol.ntuples));
  for (int64_t b = 0; b < ocol.ntuples; ++b) all_cols[b] = b;
  const bool all_pairs = old_dense || new_dense;

  std::vector<int64_t> run_of(ndim), pos_old(ndim);
  for (int l = 0; l < ndim; ++l) pos_old[l] = l < old_nrr ? l : l - old_nrr;
  std::vector<Digit> dig(ndim);
  const size_t kr = orl.size(), kc = ocl.size(), knr = nrl.size(), knc = ncl.size();
  for (int64_t a = 0; a < orow.ntuples; ++a) {
    const std::vector<int64_t> &cols = all_pairs ? all_cols : by_q[orow.q[a]];
    for (int l = 0; l < static_cast<int>(kr); ++l) run_of[l] = orow.runidx[a * kr + l];
    for (int64_t b : cols) {
      for (size_t l = 0; l < kc; ++l) run_of[old_nrr + l] = ocol.runidx[b * kc + l];
      int64_t ir = 0, ic = 0;
      for (size_t i = 0; i < knr; ++i) ir = ir * nrow.nruns[i] + run_of[nrl[i]];
      for (size_t i = 0; i < knc; ++i) ic = ic * ncol.nruns[i] + run_of[ncl[i]];
      const uint8_t qn = nrow.q[ir];
      if (!new_dense && (ncol.q[ic] != qn || !n_present[qn])) continue;  // not stored
      const uint8_t qo = orow.q[a];
      const bool readable = old_dense || (orow.q[a] == ocol.q[b] && o_present[qo]);
      int64_t src_base = -1;
      if (readable) src_base = o_off[qo] + orow.base[a] * o_rs[qo] + ocol.base[b] * o_cs[qo];
      const uint8_t qd = new_dense ? 0 : qn;
      const int64_t dst_base = n_off[qd] + nrow.base[ir] * n_cols[qd] + ncol.base[ic];
      for (int i = 0; i < ndim; ++i) {
        const int L = axes[i];
        Digit &g = dig[i];
        g.len = runs[L][run_of[L]].len;
        if (!readable) g.s = 0;  // zero tile: nothing is read
        else if (L < old_nrr) g.s = orow.stride[a * kr + pos_old[L]] * o_rs[qo];
        else g.s = ocol.stride[b * kc + pos_old[L]] * o_cs[qo];
        if (i < new_nrr) g.d = nrow.stride[ir * knr + i] * n_cols[qd];
        else g.d = ncol.stride[ic * knc + (i - new_nrr)];
        g.l = 0;
      }
      plan->boxes.emplace_back();
      make_box(dig.data(), ndim, src_base, dst_base, src_limit, std::max<int64_t>(noff, 1), plan->boxes.back());
    }
  }
  // tiles: per box nchunks x (product of the slow lengths), numbered box by box
  const size_t nb = plan->boxes.size();
  plan->tile0.resize(nb + 1);
  int64_t nt = 0;
  for (size_t i = 0; i < nb; ++i) {
    const PermBox &bx = plan->boxes[i];
    int64_t t = bx.nchunks;
    for (int k = 0; k < bx.nslow; ++k) t *= bx.slow_len[k];
    plan->tile0[i] = static_cast<int32_t>(nt);
    nt += t;
    if (nt >= (int64_t(1) << 31)) throw std::invalid_argument("permute plan: too many tiles");
  }
  plan->tile0[nb] = static_cast<int32_t>(nt);
  plan->ntiles = nt;
  plan->tile_box.resize(static_cast<size_t>(nt));
  for (size_t i = 0; i < nb; ++i)
    std::fill(plan->tile_box.begin() + plan->tile0[i], plan->tile_box.begin() + plan->tile0[i + 1],
              static_cast<int32_t>(i));
  return plan.release();
}

} // namespace fs

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
