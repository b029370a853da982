// Structural transforms of SU(2) tensors: permute, dense -> blocks and
// blocks -> dense as one planned kernel.
//
// Each task is an elementary block (one irrep per leg) holding n_in input rows
// and n_out output rows of E coefficients; the kernel computes
// out[o, :] = sum_i W[o, i] in[i, :] with every row gathered from / scattered
// to its own strided place in the block (or dense) buffers. A tile is (task,
// 64 external indices), one per lane: each lane decodes its external
// multi-index once, gathers its input column (n_in <= 32: into registers, one
// wave per tile; 32 < n_in <= 128: into LDS, four waves per tile sharing it),
// then produces the output rows 8 at a time with the transposed structural
// matrix read as wave-uniform scalar loads. The lane order follows the
// output's fastest external index, so stores coalesce.
// This replaces the per-simple-block numba loops of
// lie_group_symmetric_tensor.py:96-265 (transpose-reshape + small GEMM +
// slice) and the per-elementary-block projections of lie:856-1109.
#include <algorithm>
#include <array>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "fs_hip.h"

namespace {

constexpr int kLanes = 64;
constexpr int kOutChunk = 8;
// The device copy of each task's W^T holds n_in rounded up to kRowGroup rows
// (the extra rows zero), so the kernels read it kRowGroup rows at a time with
// every scalar load of a group in flight together.
constexpr int kRowGroup = 4;
constexpr int kMaxLdsRows = 96;  // 96 * 64 * 8 B = 48 KiB of LDS per wave

struct TaskDev {
  int64_t w_off;
  int32_t e_size, n_in, n_out, n_out_pad;
  int32_t row0, ndig, dig0, n_in_pad;  // n_in rounded up to kRowGroup (device W^T rows)
};

// In-block offsets (er * rs + ec * cs) are validated below 2^31, so row and
// column strides and the decoded external indices are 32-bit and every address
// is one 64-bit multiply-add per stride.
struct RowDev {
  int64_t base;
  int32_t rs, cs;
};

struct DigitDev {
  double inv;  // 1 / radix: division by the radix as a multiply plus one correction
  int32_t radix, in_r, in_c, out_r, out_c, pad;
};

struct TileDev {
  int32_t task, e0;
};

// Decode the external multi-index of element e into its row/column offsets on
// the input and output sides (mixed radix, last digit fastest).
__device__ inline void decode_external(const DigitDev *__restrict__ digits, const TaskDev &tk,
                                       int32_t e, int32_t &er_in, int32_t &ec_in,
                                       int32_t &er_out, int32_t &ec_out) {
  er_in = ec_in = er_out = ec_out = 0;
  int32_t rem = e;
  for (int k = tk.ndig - 1; k >= 0; --k) {
    const DigitDev d = digits[tk.dig0 + k];
    // rem / radix through the reciprocal: exact up to one unit, corrected
    int32_t q = static_cast<int32_t>(static_cast<double>(rem) * d.inv);
    int32_t r = rem - q * d.radix;
    if (r < 0) {
      --q;
      r += d.radix;
    } else if (r >= d.radix) {
      ++q;
      r -= d.radix;
    }
    rem = q;
    er_in += r * d.in_r;
    ec_in += r * d.in_c;
    er_out += r * d.out_r;
    ec_out += r * d.out_c;
  }
}

__device__ inline int64_t row_offset(const RowDev &r, int32_t er, int32_t ec) {
  return r.base + static_cast<int64_t>(er) * r.rs + static_cast<int64_t>(ec) * r.cs;
}

// Row descriptors travel one per lane: a wave reads up to 64 of them with one
// vector load (lane l holds rows[first + min(l, count - 1)]; callers pass
// count >= 1 or, for a task without input rows, first = 0, where the task's
// first output row lies, so every read stays inside the task's rows), and
// row_from_lane() broadcasts lane j's to the
// wave through readlane. The gathers that follow then wait on no memory but
// their own, instead of one scalar descriptor load per row. Must be called
// with all lanes of the wave active.
struct LaneRow {
  uint32_t base_lo, base_hi;
  int32_t rs, cs;
};

__device__ inline LaneRow load_lane_rows(const RowDev *__restrict__ rows, int first, int count,
                                         int lane) {
  const RowDev r = rows[first + min(lane, max(count - 1, 0))];
  const uint64_t b = static_cast<uint64_t>(r.base);
  return {static_cast<uint32_t>(b), static_cast<uint32_t>(b >> 32), r.rs, r.cs};
}

__device__ inline RowDev row_from_lane(const LaneRow &l, int j) {
  const uint64_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(l.base_lo), j));
  const uint64_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(l.base_hi), j));
  RowDev r;
  r.base = static_cast<int64_t>((hi << 32) | lo);
  r.rs = __builtin_amdgcn_readlane(l.rs, j);
  r.cs = __builtin_amdgcn_readlane(l.cs, j);
  return r;
}

// Register form: each lane keeps its whole input column (n_in <= kRegIn) in
// registers, loaded back to back so the gathers overlap; every output row is
// then one wave-uniform dot product. Tasks with more input rows stream their
// inputs from the cache once per chunk of kOutChunk outputs. No LDS, so the
// occupancy is set by registers only.
constexpr int kRegIn = 32;

__global__ void __launch_bounds__(kLanes)
lie_transform_reg_kernel(const TileDev *__restrict__ tiles, const TaskDev *__restrict__ tasks,
                         const RowDev *__restrict__ rows, const DigitDev *__restrict__ digits,
                         const double *__restrict__ w, const double *__restrict__ src,
                         double *__restrict__ dst) {
  const TileDev tile = tiles[blockIdx.x];
  const TaskDev tk = tasks[tile.task];
  const int lane = threadIdx.x;
  const int32_t e = tile.e0 + lane;
  const bool active = e < tk.e_size;
  int32_t er_in = 0, ec_in = 0, er_out = 0, ec_out = 0;
  if (active) decode_external(digits, tk, e, er_in, ec_in, er_out, ec_out);
  const RowDev *in_rows = rows + tk.row0;
  const RowDev *out_rows = in_rows + tk.n_in;

  if (tk.n_in <= kRegIn) {
    double x[kRegIn];
    const LaneRow lin = load_lane_rows(in_rows, 0, tk.n_in, lane);
    const LaneRow lout = load_lane_rows(out_rows, 0, tk.n_out, lane);
    // both descriptor loads complete here, before any gather is issued: later
    // readlanes then need no wait (which would also drain the gathers)
    __builtin_amdgcn_s_waitcnt(0);
#pragma unroll
    for (int i = 0; i < kRegIn; ++i) {
      x[i] = 0.0;
      if (i < tk.n_in) {
        const RowDev r = row_from_lane(lin, i);
        if (active && r.base >= 0) x[i] = src[row_offset(r, er_in, ec_in)];
      }
    }
    // outputs in chunks of 8: each input row i contributes W^T[i, o0 .. o0+7],
    // a contiguous (wave-uniform) run of the transposed structural matrix
    // (rows padded to a multiple of 8, so the run is always in bounds)
    for (int o0 = 0; o0 < tk.n_out; o0 += kOutChunk) {
      double acc[kOutChunk];
#pragma unroll
      for (int q = 0; q < kOutChunk; ++q) acc[q] = 0.0;
      const double *wt = w + tk.w_off + o0;
#pragma unroll
      for (int i0 = 0; i0 < kRegIn; i0 += kRowGroup) {
        if (i0 < tk.n_in) {  // rows n_in .. n_in_pad - 1: zero W, zero x
#pragma unroll
          for (int j = 0; j < kRowGroup; ++j) {
            const double *wi = wt + static_cast<int64_t>(i0 + j) * tk.n_out_pad;
#pragma unroll
            for (int q = 0; q < kOutChunk; ++q) acc[q] = fma(wi[q], x[i0 + j], acc[q]);
          }
        }
      }
#pragma unroll
      for (int q = 0; q < kOutChunk; ++q) {
        if (o0 + q < tk.n_out) {
          // output rows past the first 64 (rare) come straight from the table
          const RowDev ro = o0 < kLanes ? row_from_lane(lout, o0 + q) : out_rows[o0 + q];
          if (active) dst[row_offset(ro, er_out, ec_out)] = acc[q];
        }
      }
    }
    return;
  }

  constexpr int kChunk = 2 * kOutChunk;
  for (int o0 = 0; o0 < tk.n_out; o0 += kChunk) {
    double acc[kChunk];
#pragma unroll
    for (int r = 0; r < kChunk; ++r) acc[r] = 0.0;
    const double *wt = w + tk.w_off + o0;
    const int nr = min(kChunk, tk.n_out_pad - o0);  // W rows are padded to a multiple of 8
    for (int i = 0; i < tk.n_in; ++i) {
      const RowDev r = in_rows[i];
      const double xv = (active && r.base >= 0) ? src[row_offset(r, er_in, ec_in)] : 0.0;
      const double *wi = wt + static_cast<int64_t>(i) * tk.n_out_pad;
#pragma unroll
      for (int q = 0; q < kChunk; ++q)
        if (q < nr) acc[q] = fma(wi[q], xv, acc[q]);
    }
    if (active) {
      const int nw = min(kChunk, tk.n_out - o0);
      for (int q = 0; q < nw; ++q) {
        const RowDev ro = out_rows[o0 + q];
        dst[row_offset(ro, er_out, ec_out)] = acc[q];
      }
    }
  }
}

// Wide form for tasks with kRegIn < n_in <= kWideMaxIn: four waves share one
// tile of 64 external indices. The waves gather the input column into LDS,
// kWideBatch rows per wave with every load of a batch in flight at once, then
// each wave produces every fourth chunk of 8 output rows from LDS. Against the
// streamed path of the register form (one dependent gather per input row and
// output chunk) this turns ~n_in * n_out / 16 serial memory latencies per wave
// into ~n_in / 32.
constexpr int kWideWaves = 4;
constexpr int kWideMaxIn = 128;  // 128 * 64 * 8 B = 64 KiB of LDS per workgroup
constexpr int kWideBatch = 8;

__global__ void __launch_bounds__(kLanes * kWideWaves)
lie_transform_wide_kernel(const TileDev *__restrict__ tiles, const TaskDev *__restrict__ tasks,
                          const RowDev *__restrict__ rows, const DigitDev *__restrict__ digits,
                          const double *__restrict__ w, const double *__restrict__ src,
                          double *__restrict__ dst) {
  extern __shared__ double xs[];  // [n_in][kLanes]
  const TileDev tile = tiles[blockIdx.x];
  const TaskDev tk = tasks[tile.task];
  const int lane = threadIdx.x % kLanes;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kLanes);
  const int32_t e = tile.e0 + lane;
  const bool active = e < tk.e_size;
  int32_t er_in = 0, ec_in = 0, er_out = 0, ec_out = 0;
  if (active) decode_external(digits, tk, e, er_in, ec_in, er_out, ec_out);
  const RowDev *in_rows = rows + tk.row0;
  const RowDev *out_rows = in_rows + tk.n_in;

  // rows 0..63 and 64..127 of the input column, one descriptor per lane
  // (n_in > kRegIn >= 1 here; with n_in <= 64 the second read repeats row n_in - 1)
  const LaneRow lin0 = load_lane_rows(in_rows, 0, tk.n_in, lane);
  const LaneRow lin1 = load_lane_rows(in_rows, min(kLanes, tk.n_in - 1), max(tk.n_in - kLanes, 1), lane);
  for (int i0 = wave * kWideBatch; i0 < tk.n_in; i0 += kWideWaves * kWideBatch) {
    double v[kWideBatch];
    const LaneRow &lr = i0 < kLanes ? lin0 : lin1;  // a batch never straddles row 64
#pragma unroll
    for (int j = 0; j < kWideBatch; ++j) {
      v[j] = 0.0;
      if (i0 + j < tk.n_in) {
        const RowDev r = row_from_lane(lr, (i0 + j) % kLanes);
        if (active && r.base >= 0) v[j] = src[row_offset(r, er_in, ec_in)];
      }
    }
#pragma unroll
    for (int j = 0; j < kWideBatch; ++j)
      if (i0 + j < tk.n_in_pad) xs[(i0 + j) * kLanes + lane] = v[j];  // zero past n_in
  }
  __syncthreads();

  for (int o0 = wave * kOutChunk; o0 < tk.n_out; o0 += kWideWaves * kOutChunk) {
    const LaneRow lout = load_lane_rows(out_rows, o0, tk.n_out - o0, lane);
    double acc[kOutChunk];
#pragma unroll
    for (int q = 0; q < kOutChunk; ++q) acc[q] = 0.0;
    const double *wt = w + tk.w_off + o0;  // W^T rows padded to a multiple of 8
    for (int i0 = 0; i0 < tk.n_in_pad; i0 += kRowGroup) {
#pragma unroll
      for (int j = 0; j < kRowGroup; ++j) {
        const double x = xs[(i0 + j) * kLanes + lane];
        const double *wi = wt + static_cast<int64_t>(i0 + j) * tk.n_out_pad;
#pragma unroll
        for (int q = 0; q < kOutChunk; ++q) acc[q] = fma(wi[q], x, acc[q]);
      }
    }
#pragma unroll
    for (int q = 0; q < kOutChunk; ++q) {
      if (o0 + q < tk.n_out) {
        const RowDev ro = row_from_lane(lout, q);
        if (active) dst[row_offset(ro, er_out, ec_out)] = acc[q];
      }
    }
  }
}

__global__ void __launch_bounds__(kLanes)
lie_transform_kernel(const TileDev *__restrict__ tiles, const TaskDev *__restrict__ tasks,
                     const RowDev *__restrict__ rows, const DigitDev *__restrict__ digits,
                     const double *__restrict__ w, const double *__restrict__ src,
                     double *__restrict__ dst, int lds_rows) {
  extern __shared__ double xs[];
  const TileDev tile = tiles[blockIdx.x];
  const TaskDev tk = tasks[tile.task];
  const int lane = threadIdx.x;
  const int32_t e = tile.e0 + lane;
  const bool active = e < tk.e_size;

  int32_t er_in = 0, ec_in = 0, er_out = 0, ec_out = 0;
  if (active) decode_external(digits, tk, e, er_in, ec_in, er_out, ec_out);

  const RowDev *in_rows = rows + tk.row0;
  const RowDev *out_rows = in_rows + tk.n_in;
  const bool staged = tk.n_in <= lds_rows;
  if (staged) {
    for (int i = 0; i < tk.n_in; ++i) {
      const RowDev r = in_rows[i];
      double v = 0.0;
      if (active && r.base >= 0) v = src[row_offset(r, er_in, ec_in)];
      xs[i * kLanes + lane] = v;
    }
  }

  for (int o0 = 0; o0 < tk.n_out; o0 += kOutChunk) {
    double acc[kOutChunk];
#pragma unroll
    for (int r = 0; r < kOutChunk; ++r) acc[r] = 0.0;
    const double *wt = w + tk.w_off + o0;
    for (int i = 0; i < tk.n_in; ++i) {
      double x;
      if (staged) {
        x = xs[i * kLanes + lane];
      } else {
        const RowDev r = in_rows[i];
        x = (active && r.base >= 0) ? src[row_offset(r, er_in, ec_in)] : 0.0;
      }
      const double *wi = wt + static_cast<int64_t>(i) * tk.n_out_pad;
#pragma unroll
      for (int r = 0; r < kOutChunk; ++r) acc[r] = fma(wi[r], x, acc[r]);
    }
    if (active) {
      const int nr = min(kOutChunk, tk.n_out - o0);
      for (int r = 0; r < nr; ++r) {
        const RowDev ro = out_rows[o0 + r];
        dst[row_offset(ro, er_out, ec_out)] = acc[r];
      }
    }
  }
}

} // namespace

// Structural matrices of one permute pattern on the device, shared by every
// plan built from that pattern (they do not depend on the degeneracies):
// W^T of each task padded to kRowGroup rows, tasks with the same matrix
// sharing one copy.
struct fs_lie_weights {
  int device = -1;
  mutable fs::DeviceTables tables;  // prepare() from the plans' launches
  int64_t nw = 0;                // length of the unpadded W it was built from
  std::vector<int64_t> off;      // per pattern task: offset of its padded copy
  // per pattern task: (w_off, n_in, n_out_pad) the copy was packed with; a
  // plan may use these weights only if its tasks agree on all three
  std::vector<std::array<int64_t, 3>> layout;
  uint64_t fingerprint = 0;  // pattern_fingerprint of the pattern it was built from
};

struct fs_lie_plan {
  int device = -1;
  mutable fs::DeviceTables tables;  // tiles, tasks, rows, digits, W^T (256-byte aligned parts)
  const fs_lie_weights *weights = nullptr;  // shared W^T (then not in `tables`)
  const TileDev *tiles = nullptr;
  const TaskDev *tasks = nullptr;
  const RowDev *rows = nullptr;
  const DigitDev *digits = nullptr;
  const double *w = nullptr;
  int64_t ntiles = 0;       // all tiles: register-form tiles, then wide-form tiles
  int64_t ntiles_wide = 0;  // tiles of tasks with kRegIn < n_in <= kWideMaxIn
  int32_t lds_rows = 0;
  int32_t wide_rows = 0;    // largest n_in_pad among the wide-form tasks
  int64_t src_size = 0, dst_size = 0;
  double flops = 0.0, coeffs = 0.0;
};

namespace {

int check_task(const fs_lie_task &t, int64_t nrows, const fs_lie_row *rows, int64_t ndigits,
               const fs_lie_digit *digits, int64_t nw, int64_t src_size, int64_t dst_size) {
  FS_CHECK_ARG(t.e_size >= 1 && t.e_size < (int64_t(1) << 31), "lie plan: bad external size");
  FS_CHECK_ARG(t.n_in >= 0 && t.n_out >= 0 && t.n_in < (1 << 24) && t.n_out < (1 << 24),
               "lie plan: bad row counts");
  FS_CHECK_ARG(t.n_out_pad >= t.n_out && t.n_out_pad % kOutChunk == 0,
               "lie plan: n_out_pad must be a multiple of 8 >= n_out");
  FS_CHECK_ARG(t.ndig >= 1 && t.ndig <= FS_LIE_MAX_LEGS, "lie plan: bad digit count");
  FS_CHECK_ARG(t.dig0 >= 0 && t.dig0 + t.ndig <= ndigits, "lie plan: digits out of range");
  FS_CHECK_ARG(t.row0 >= 0 && t.row0 + t.n_in + t.n_out <= nrows, "lie plan: rows out of range");
  FS_CHECK_ARG(t.w_off >= 0 && t.w_off + t.n_in * t.n_out_pad <= nw, "lie plan: W out of range");
  int64_t prod = 1, mr_in = 0, mc_in = 0, mr_out = 0, mc_out = 0;
  for (int64_t k = 0; k < t.ndig; ++k) {
    const fs_lie_digit &d = digits[t.dig0 + k];
    FS_CHECK_ARG(d.radix >= 1 && d.in_r >= 0 && d.in_c >= 0 && d.out_r >= 0 && d.out_c >= 0,
                 "lie plan: bad digit");
    FS_CHECK_ARG(d.in_r < (int64_t(1) << 31) && d.in_c < (int64_t(1) << 31) &&
                     d.out_r < (int64_t(1) << 31) && d.out_c < (int64_t(1) << 31),
                 "lie plan: digit stride too large");
    prod *= d.radix;
    FS_CHECK_ARG(prod <= t.e_size, "lie plan: digits exceed external size");
    mr_in += (d.radix - 1) * d.in_r;
    mc_in += (d.radix - 1) * d.in_c;
    mr_out += (d.radix - 1) * d.out_r;
    mc_out += (d.radix - 1) * d.out_c;
    FS_CHECK_ARG(std::max(std::max(mr_in, mc_in), std::max(mr_out, mc_out)) < (int64_t(1) << 62),
                 "lie plan: external index too large");
  }
  FS_CHECK_ARG(prod == t.e_size, "lie plan: digit radices do not multiply to the external size");
  FS_CHECK_ARG(std::max(std::max(mr_in, mc_in), std::max(mr_out, mc_out)) < (int64_t(1) << 31),
               "lie plan: external index beyond 32 bits");
  for (int64_t i = 0; i < t.n_in + t.n_out; ++i) {
    const fs_lie_row &r = rows[t.row0 + i];
    const bool input = i < t.n_in;
    if (input && r.base < 0) continue;  // absent input block
    FS_CHECK_ARG(r.base >= 0 && r.rs >= 0 && r.cs >= 0, "lie plan: bad row entry");
    FS_CHECK_ARG(r.rs < (int64_t(1) << 31) && r.cs < (int64_t(1) << 31), "lie plan: stride beyond 32 bits");
    int64_t a = 0, b = 0, last = 0;
    const bool ovf = __builtin_mul_overflow(input ? mr_in : mr_out, r.rs, &a) ||
                     __builtin_mul_overflow(input ? mc_in : mc_out, r.cs, &b) ||
                     __builtin_add_overflow(r.base, a, &last) ||
                     __builtin_add_overflow(last, b, &last);
    FS_CHECK_ARG(!ovf && last < (input ? src_size : dst_size),
                 "lie plan: row entry %lld out of bounds", static_cast<long long>(t.row0 + i));
  }
  return FS_OK;
}

// Padded device copy of W (see fs_lie_weights) for tasks given by their
// (w_off, n_in, n_out_pad); off[t] receives task t's offset in wdev.
template <class Get>
void repack_w(int64_t ntasks, Get get, const double *w, std::vector<double> &wdev,
              std::vector<int64_t> &off) {
  std::map<std::tuple<int64_t, int64_t, int64_t>, int64_t> copies;
  off.resize(static_cast<size_t>(ntasks));
  for (int64_t t = 0; t < ntasks; ++t) {
    int64_t w_off, n_in, n_out_pad;
    get(t, w_off, n_in, n_out_pad);
    const int64_t n_in_pad = (n_in + kRowGroup - 1) / kRowGroup * kRowGroup;
    const auto key = std::make_tuple(w_off, n_in, n_out_pad);
    auto found = copies.find(key);
    if (found == copies.end()) {
      const int64_t o = static_cast<int64_t>(wdev.size());
      if (n_in * n_out_pad > 0) wdev.insert(wdev.end(), w + w_off, w + w_off + n_in * n_out_pad);
      wdev.resize(static_cast<size_t>(o + n_in_pad * n_out_pad), 0.0);
      found = copies.emplace(key, o).first;
    }
    off[t] = found->second;
  }
}

// FNV-1a over the pattern's shape and per-task arrays (legs, permutation,
// irreps counts, task irreps indices and row counts): weights built from one
// pattern are refused for a plan of another.
bool pattern_arrays_present(const fs_lie_pattern &p) {
  if (p.ndim < 1 || p.ntasks < 0) return false;
  if (!p.axes || !p.nirr) return false;
  if (p.ntasks == 0) return true;
  for (const int64_t *a : {p.sig_idx, p.ir_in, p.ic_in, p.ir_out, p.ic_out, p.n_in, p.n_out, p.w_off,
                           p.n_out_pad})
    if (!a) return false;
  return true;
}

// Caller checks pattern_arrays_present first.
uint64_t pattern_fingerprint(const fs_lie_pattern &p) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&h](int64_t v) {
    for (int b = 0; b < 8; ++b) {
      h ^= static_cast<uint64_t>(v >> (8 * b)) & 0xffu;
      h *= 1099511628211ull;
    }
  };
  auto mix_arr = [&mix](const int64_t *a, int64_t n) {
    for (int64_t i = 0; i < n; ++i) mix(a[i]);
  };
  for (int64_t v : {p.ndim, p.nrr, p.nrr_out, p.ntasks, p.nt_in, p.nt_out}) mix(v);
  mix_arr(p.axes, p.ndim);
  mix_arr(p.nirr, p.ndim);
  mix_arr(p.sig_idx, p.ntasks * p.ndim);
  for (const int64_t *a : {p.ir_in, p.ic_in, p.ir_out, p.ic_out, p.n_in, p.n_out, p.w_off, p.n_out_pad})
    mix_arr(a, p.ntasks);
  return h;
}

// Plan from tables; W either uploaded with the plan (shared == nullptr) or
// taken from a shared weights object built from the same pattern.
int create_impl(int64_t ntasks, const fs_lie_task *tasks, int64_t nrows, const fs_lie_row *rows,
                int64_t ndigits, const fs_lie_digit *digits, int64_t nw, const double *w,
                const fs_lie_weights *shared, int64_t src_size, int64_t dst_size,
                fs_lie_plan **plan) {
  {
    FS_CHECK_ARG(plan != nullptr, "lie plan: null output");
    *plan = nullptr;
    if (shared) {
      FS_CHECK_ARG(static_cast<int64_t>(shared->off.size()) == ntasks, "lie plan: weights of another pattern");
      FS_CHECK_ARG(ntasks == 0 || tasks, "lie plan: null table");
      for (int64_t t = 0; t < ntasks; ++t) {
        const std::array<int64_t, 3> &l = shared->layout[static_cast<size_t>(t)];
        FS_CHECK_ARG(l[0] == tasks[t].w_off && l[1] == tasks[t].n_in && l[2] == tasks[t].n_out_pad,
                     "lie plan: weights of another pattern (task %lld layout differs)",
                     static_cast<long long>(t));
      }
      nw = shared->nw;
    }
    FS_CHECK_ARG(ntasks >= 0 && nrows >= 0 && ndigits >= 0 && nw >= 0, "lie plan: negative size");
    FS_CHECK_ARG(ntasks == 0 || (tasks && rows && digits), "lie plan: null table");
    FS_CHECK_ARG(shared || nw == 0 || w, "lie plan: null W");
    FS_CHECK_ARG(src_size >= 0 && dst_size >= 0, "lie plan: negative buffer size");
    FS_CHECK_ARG(nrows < (int64_t(1) << 31) && ndigits < (int64_t(1) << 31),
                 "lie plan: tables too large");
    auto p = std::make_unique<fs_lie_plan>();
    std::vector<TaskDev> tdev;
    std::vector<int32_t> reg_ids, wide_ids;
    for (int64_t t = 0; t < ntasks; ++t) {
      int st = check_task(tasks[t], nrows, rows, ndigits, digits, nw, src_size, dst_size);
      if (st != FS_OK) return st;
    }
    // device copy of W^T with every task's rows padded to a multiple of
    // kRowGroup (tasks sharing a structural matrix share its copy), unless
    // the pattern's shared copy is used
    std::vector<double> wdev;
    std::vector<int64_t> w_off_own;
    if (!shared)
      repack_w(ntasks, [&](int64_t t, int64_t &wo, int64_t &ni, int64_t &nop) {
        wo = tasks[t].w_off; ni = tasks[t].n_in; nop = tasks[t].n_out_pad;
      }, w, wdev, w_off_own);
    const std::vector<int64_t> &w_offs = shared ? shared->off : w_off_own;
    tdev.reserve(static_cast<size_t>(ntasks));
    int64_t max_in = 0, max_wide_in = 0;
    for (int64_t t = 0; t < ntasks; ++t) {
      const fs_lie_task &s = tasks[t];
      TaskDev d;
      const int64_t n_in_pad = (s.n_in + kRowGroup - 1) / kRowGroup * kRowGroup;
      d.w_off = w_offs[t];
      d.e_size = static_cast<int32_t>(s.e_size);
      d.n_in = static_cast<int32_t>(s.n_in);
      d.n_out = static_cast<int32_t>(s.n_out);
      d.n_out_pad = static_cast<int32_t>(s.n_out_pad);
      d.row0 = static_cast<int32_t>(s.row0);
      d.ndig = static_cast<int32_t>(s.ndig);
      d.dig0 = static_cast<int32_t>(s.dig0);
      d.n_in_pad = static_cast<int32_t>(n_in_pad);
      const int32_t id = static_cast<int32_t>(tdev.size());
      tdev.push_back(d);
      if (s.n_out == 0) continue;
      max_in = std::max(max_in, s.n_in);
      if (s.n_in > kRegIn && s.n_in <= kWideMaxIn) {
        wide_ids.push_back(id);
        max_wide_in = std::max(max_wide_in, n_in_pad);
      } else {
        reg_ids.push_back(id);
      }
      p->flops += 2.0 * double(s.n_in) * double(s.n_out) * double(s.e_size);
      p->coeffs += double(s.n_in + s.n_out) * double(s.e_size);
    }
    // register-form tasks heaviest first (n_in * n_out per element), so that
    // the longest waves start early instead of forming the launch's tail
    std::stable_sort(reg_ids.begin(), reg_ids.end(), [&](int32_t a, int32_t b) {
      return int64_t(tdev[a].n_in) * tdev[a].n_out > int64_t(tdev[b].n_in) * tdev[b].n_out;
    });
    // one image, 256-byte aligned parts (tiles, tasks, rows, digits, W^T),
    // written straight into the tables' staging memory
    size_t n_tiles = 0, n_reg_tiles = 0;
    for (int32_t id : reg_ids) n_reg_tiles += static_cast<size_t>((tdev[id].e_size + kLanes - 1) / kLanes);
    n_tiles = n_reg_tiles;
    for (int32_t id : wide_ids) n_tiles += static_cast<size_t>((tdev[id].e_size + kLanes - 1) / kLanes);
    FS_CHECK_ARG(n_tiles < (size_t(1) << 31), "lie plan: too many tiles");
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t b_tiles = n_tiles * sizeof(TileDev), b_tasks = tdev.size() * sizeof(TaskDev);
    const size_t b_rows = static_cast<size_t>(nrows) * sizeof(RowDev);
    const size_t b_dig = static_cast<size_t>(ndigits) * sizeof(DigitDev);
    const size_t b_w = wdev.size() * sizeof(double);
    const size_t o_tasks = al(b_tiles), o_rows = o_tasks + al(b_tasks), o_dig = o_rows + al(b_rows);
    const size_t o_w = o_dig + al(b_dig), total = o_w + al(b_w) + 256;
    FS_HIP_CHECK(hipGetDevice(&p->device));
    if (shared)
      FS_CHECK_ARG(shared->device == p->device, "lie plan: weights on device %d, plan on device %d",
                   shared->device, p->device);
    char *img = nullptr;
    int st = p->tables.begin(total, &img);
    if (st != FS_OK) return st;
    TileDev *tp = reinterpret_cast<TileDev *>(img);
    for (const std::vector<int32_t> *ids : {&reg_ids, &wide_ids})
      for (int32_t id : *ids)
        for (int32_t e0 = 0; e0 < tdev[id].e_size; e0 += kLanes) *tp++ = {id, e0};
    if (b_tasks) std::memcpy(img + o_tasks, tdev.data(), b_tasks);
    RowDev *rp = reinterpret_cast<RowDev *>(img + o_rows);
    for (int64_t i = 0; i < nrows; ++i)
      rp[i] = {rows[i].base, static_cast<int32_t>(rows[i].rs), static_cast<int32_t>(rows[i].cs)};
    DigitDev *dp = reinterpret_cast<DigitDev *>(img + o_dig);
    for (int64_t i = 0; i < ndigits; ++i) {
      const fs_lie_digit &d = digits[i];
      dp[i] = {1.0 / static_cast<double>(d.radix), static_cast<int32_t>(d.radix),
               static_cast<int32_t>(d.in_r), static_cast<int32_t>(d.in_c),
               static_cast<int32_t>(d.out_r), static_cast<int32_t>(d.out_c), 0};
    }
    if (b_w) std::memcpy(img + o_w, wdev.data(), b_w);
    // the alignment gaps and the tail stay zero, as in a zero-filled image
    const size_t part_end[5] = {b_tiles, o_tasks + b_tasks, o_rows + b_rows, o_dig + b_dig, o_w + b_w};
    const size_t next[5] = {o_tasks, o_rows, o_dig, o_w, total};
    for (int q = 0; q < 5; ++q) std::memset(img + part_end[q], 0, next[q] - part_end[q]);
    st = p->tables.commit();
    if (st != FS_OK) return st;
    const char *mem = p->tables.data();
    const char *d_tiles = b_tiles ? mem : nullptr;
    const char *d_tasks = b_tasks ? mem + o_tasks : nullptr;
    const char *d_rows = b_rows ? mem + o_rows : nullptr;
    const char *d_dig = b_dig ? mem + o_dig : nullptr;
    const char *d_w = b_w ? mem + o_w : nullptr;
    p->tiles = reinterpret_cast<const TileDev *>(d_tiles);
    p->tasks = reinterpret_cast<const TaskDev *>(d_tasks);
    p->rows = reinterpret_cast<const RowDev *>(d_rows);
    p->digits = reinterpret_cast<const DigitDev *>(d_dig);
    p->w = shared ? reinterpret_cast<const double *>(shared->tables.data())
                  : reinterpret_cast<const double *>(d_w);
    p->weights = shared;
    p->ntiles = static_cast<int64_t>(n_tiles);
    p->ntiles_wide = static_cast<int64_t>(n_tiles - n_reg_tiles);
    p->wide_rows = static_cast<int32_t>(max_wide_in);
    p->lds_rows = static_cast<int32_t>(std::min<int64_t>(max_in, kMaxLdsRows));
    p->src_size = src_size;
    p->dst_size = dst_size;
    *plan = p.release();
    return FS_OK;
  }
}

} // namespace

extern "C" {

int fs_lie_plan_create(int64_t ntasks, const fs_lie_task *tasks, int64_t nrows,
                       const fs_lie_row *rows, int64_t ndigits, const fs_lie_digit *digits,
                       int64_t nw, const double *w, int64_t src_size, int64_t dst_size,
                       fs_lie_plan **plan) {
  return fs::guarded([&]() -> int {
    return create_impl(ntasks, tasks, nrows, rows, ndigits, digits, nw, w, nullptr, src_size,
                       dst_size, plan);
  });
}

int fs_lie_weights_create(const fs_lie_pattern *pattern, int64_t nw, const double *w,
                          fs_lie_weights **weights) {
  return fs::guarded([&]() -> int {
    FS_CHECK_ARG(pattern && weights, "lie weights: null argument");
    *weights = nullptr;
    FS_CHECK_ARG(pattern_arrays_present(*pattern), "lie weights: pattern array missing");
    const int64_t ns = pattern->ntasks;
    FS_CHECK_ARG(ns >= 0 && nw >= 0 && (nw == 0 || w), "lie weights: bad W");
    for (int64_t t = 0; t < ns; ++t) {
      const int64_t wo = pattern->w_off[t], ni = pattern->n_in[t], nop = pattern->n_out_pad[t];
      FS_CHECK_ARG(wo >= 0 && ni >= 0 && nop >= 0 && nop % kOutChunk == 0 && wo + ni * nop <= nw,
                   "lie weights: task %lld out of W", static_cast<long long>(t));
    }
    auto q = std::make_unique<fs_lie_weights>();
    std::vector<double> wdev;
    repack_w(ns, [&](int64_t t, int64_t &wo, int64_t &ni, int64_t &nop) {
      wo = pattern->w_off[t]; ni = pattern->n_in[t]; nop = pattern->n_out_pad[t];
    }, w, wdev, q->off);
    q->fingerprint = pattern_fingerprint(*pattern);
    q->layout.resize(static_cast<size_t>(ns));
    for (int64_t t = 0; t < ns; ++t)
      q->layout[static_cast<size_t>(t)] = {pattern->w_off[t], pattern->n_in[t], pattern->n_out_pad[t]};
    FS_HIP_CHECK(hipGetDevice(&q->device));
    q->nw = nw;
    int st = q->tables.upload(wdev.data(), wdev.size() * sizeof(double));
    if (st != FS_OK) return st;
    *weights = q.release();
    return FS_OK;
  });
}

int fs_lie_weights_destroy(fs_lie_weights *weights) {
  return fs::guarded([&]() -> int {
    if (!weights) return FS_OK;
    const int st = weights->tables.release();
    delete weights;
    return st;
  });
}

int fs_lie_plan_check(int64_t ntasks, const fs_lie_task *tasks, int64_t nrows,
                      const fs_lie_row *rows, int64_t ndigits, const fs_lie_digit *digits,
                      int64_t nw, int64_t src_size, int64_t dst_size) {
  return fs::guarded([&]() -> int {
    FS_CHECK_ARG(ntasks >= 0 && nrows >= 0 && ndigits >= 0 && nw >= 0, "lie plan: negative size");
    FS_CHECK_ARG(ntasks == 0 || (tasks && rows && digits), "lie plan: null table");
    for (int64_t t = 0; t < ntasks; ++t) {
      int st = check_task(tasks[t], nrows, rows, ndigits, digits, nw, src_size, dst_size);
      if (st != FS_OK) return st;
    }
    return FS_OK;
  });
}

int fs_lie_plan_info(const fs_lie_plan *plan, int64_t *ntiles, double *flops, double *coeffs) {
  return fs::guarded([&]() -> int {
    FS_CHECK_ARG(plan != nullptr, "lie plan: null plan");
    if (ntiles) *ntiles = plan->ntiles;
    if (flops) *flops = plan->flops;
    if (coeffs) *coeffs = plan->coeffs;
    return FS_OK;
  });
}

int fs_lie_plan_execute_f64(const fs_lie_plan *plan, const double *src, double *dst,
                            void *stream) {
  return fs::guarded([&]() -> int {
    FS_CHECK_ARG(plan != nullptr, "lie plan: null plan");
    if (plan->ntiles == 0) return FS_OK;
    FS_CHECK_ARG(dst != nullptr, "lie plan: null destination");
    FS_CHECK_ARG(src != nullptr || plan->src_size == 0, "lie plan: null source");
    int dev = 0;
    FS_HIP_CHECK(hipGetDevice(&dev));
    FS_CHECK_ARG(dev == plan->device, "lie plan was built on device %d, current device is %d",
                 plan->device, dev);
    static const int mode = [] {
      // tuning only: 1 = one-wave LDS-staged form for every task, 2 = register
      // form for every task (wide tasks on its streamed path)
      const char *e = std::getenv("FS_LIE_KERNEL");
      return e ? std::atoi(e) : 0;
    }();
    int st = plan->tables.prepare(fs::as_stream(stream));
    if (st == FS_OK && plan->weights) st = plan->weights->tables.prepare(fs::as_stream(stream));
    if (st != FS_OK) return st;
    if (mode == 1) {
      const size_t lds = static_cast<size_t>(plan->lds_rows) * kLanes * sizeof(double);
      hipLaunchKernelGGL(lie_transform_kernel, dim3(static_cast<unsigned>(plan->ntiles)),
                         dim3(kLanes), lds, fs::as_stream(stream), plan->tiles, plan->tasks,
                         plan->rows, plan->digits, plan->w, src, dst, plan->lds_rows);
      FS_LAUNCH_CHECK("lie_transform_kernel");
    } else {
      const int64_t n_wide = mode == 2 ? 0 : plan->ntiles_wide;
      const int64_t n_reg = plan->ntiles - n_wide;
      if (n_wide > 0) {
        // every wide task has n_in <= wide_rows <= kWideMaxIn: the LDS column fits
        const size_t lds = static_cast<size_t>(plan->wide_rows) * kLanes * sizeof(double);
        hipLaunchKernelGGL(lie_transform_wide_kernel, dim3(static_cast<unsigned>(n_wide)),
                           dim3(kLanes * kWideWaves), lds, fs::as_stream(stream),
                           plan->tiles + n_reg, plan->tasks, plan->rows, plan->digits, plan->w,
                           src, dst);
        FS_LAUNCH_CHECK("lie_transform_wide_kernel");
      }
      if (n_reg > 0) {
        hipLaunchKernelGGL(lie_transform_reg_kernel, dim3(static_cast<unsigned>(n_reg)),
                           dim3(kLanes), 0, fs::as_stream(stream), plan->tiles, plan->tasks,
                           plan->rows, plan->digits, plan->w, src, dst);
        FS_LAUNCH_CHECK("lie_transform_reg_kernel");
      }
    }
    return FS_OK;
  });
}

int fs_lie_plan_destroy(fs_lie_plan *plan) {
  return fs::guarded([&]() -> int {
    if (!plan) return FS_OK;
    const int st = plan->tables.release();
    delete plan;
    return st;
  });
}

int fs_lie_plan_build(const fs_lie_pattern *pattern, int64_t ndegen, const int64_t *degen,
                      const int64_t *in_block, int64_t nblocks, const int64_t *block_off,
                      const int64_t *block_rs, const int64_t *block_cs,
                      const fs_lie_weights *weights, int64_t src_size, int64_t *shapes,
                      fs_lie_plan **plan) {
  return fs::guarded([&]() -> int {
    FS_CHECK_ARG(pattern && plan && shapes && weights, "lie plan build: null argument");
    *plan = nullptr;
    const int64_t ns = pattern->ntasks, nd = pattern->ndim;
    FS_CHECK_ARG(ns >= 0 && nd >= 1 && nd <= FS_LIE_MAX_LEGS, "lie plan build: bad pattern");
    FS_CHECK_ARG(pattern_arrays_present(*pattern), "lie plan build: pattern array missing");
    FS_CHECK_ARG(pattern_fingerprint(*pattern) == weights->fingerprint,
                 "lie plan build: weights of another pattern");
    int64_t nrows = 0;
    for (int64_t t = 0; t < ns; ++t) {
      FS_CHECK_ARG(pattern->n_in[t] >= 0 && pattern->n_out[t] >= 0, "lie plan build: negative row count");
      nrows += pattern->n_in[t] + pattern->n_out[t];
    }
    std::vector<int64_t> tasks(static_cast<size_t>(ns) * 8), rows(static_cast<size_t>(nrows) * 3);
    std::vector<int64_t> digits(static_cast<size_t>(ns * nd) * 5);
    int st = fs_lie_plan_tables(pattern, ndegen, degen, in_block, nblocks, block_off, block_rs,
                                block_cs, tasks.data(), nrows, rows.data(), digits.data(), shapes);
    if (st != FS_OK) return st;
    int64_t dst_size = 0;
    for (int64_t t = 0; t < pattern->nt_out; ++t) dst_size += shapes[2 * t] * shapes[2 * t + 1];
    static_assert(sizeof(fs_lie_task) == 8 * sizeof(int64_t) && sizeof(fs_lie_row) == 3 * sizeof(int64_t) &&
                      sizeof(fs_lie_digit) == 5 * sizeof(int64_t),
                  "table layouts");
    return create_impl(ns, reinterpret_cast<const fs_lie_task *>(tasks.data()), nrows,
                       reinterpret_cast<const fs_lie_row *>(rows.data()), ns * nd,
                       reinterpret_cast<const fs_lie_digit *>(digits.data()), 0, nullptr, weights,
                       src_size, dst_size, plan);
  });
}

} // extern "C"
