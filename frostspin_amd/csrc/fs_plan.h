// Abelian permute plan: a list of boxes built by the planner on the host,
// mirrored to device memory on first execution; the executor cuts every box
// into tiles itself.
//
// A box is the set of elements whose index on every leg lies in one run of
// equal charges of that leg (a "run tuple"). Inside it the position in the old
// block and in the new block are both affine in the legs' digits, so the
// permute of a box is a dense strided copy. A box larger than one LDS tile is
// cut along its slowest destination digits: the slowest ones (the "slow"
// digits) contribute one index per tile, the next one (the "cut" digit) is cut
// into chunks, the faster ones are whole in every tile. All tiles of a box
// then share one digit layout; only the last chunk of the cut digit can be
// shorter, which changes one length of the source-order digit list (the
// destination-order list holds the cut digit in its last entry, whose length
// never enters the digit decoding). A plan is therefore O(#boxes), not
// O(#tiles): the record of a tile is derived on the device from its box.
#pragma once

#include <cstdint>
#include <mutex>
#include <vector>

#include "frostspin_amd.h"

// Elements of one tile (LDS staging buffer of the permute kernel).
constexpr int64_t kPermuteTileVol = 4096;
// Digits (legs with more than one state in the box) of one tile.
constexpr int kPermuteMaxDigits = FS_PERMUTE_MAX_DIGITS;

// One box. Source digits fastest first (flat tile index f in source order:
// digit k is (f / product of the faster lengths) mod len), with staging
// offsets in destination order; destination digits fastest first likewise.
// smag / dmag: ceil(2^32 / len), so that q = umulhi(f, mag) = f / len for the
// flat indices f < kPermuteTileVol of a tile.
struct PermBox {
  int64_t src_base, dst_base;  // first element; src_base < 0: absent old block, the box is zeros
  int64_t cut_s, cut_d;        // source / destination stride of the cut digit
  int64_t slow_s[kPermuteMaxDigits], slow_d[kPermuteMaxDigits];  // slow digits, fastest first
  int32_t slow_len[kPermuteMaxDigits];
  int32_t nslow;
  int32_t nchunks, chunk;  // chunks of the cut digit and their length (1, 0: nothing is cut)
  int32_t vol, vol_edge;   // elements of a tile, of a tile in the last chunk
  int32_t ns, nd;
  // last chunk: source entry s_cut (-1: no change) gets length s_len_edge
  // (magic s_mag_edge), or is dropped when that length is 1
  int32_t s_cut, s_len_edge, s_drop_edge;
  uint32_t s_mag_edge;
  int32_t d_len_edge;  // last chunk: length of the last destination entry
  int32_t pad_;
  uint32_t smag[kPermuteMaxDigits];
  int32_t slen[kPermuteMaxDigits], sstr[kPermuteMaxDigits], lstr[kPermuteMaxDigits];
  uint32_t dmag[kPermuteMaxDigits];
  int32_t dlen[kPermuteMaxDigits], dstr[kPermuteMaxDigits];
};

struct fs_permute_plan {
  // new side (per new block)
  std::vector<int8_t> new_irreps;
  std::vector<int64_t> new_shapes;  // 2 per block
  int64_t new_ncoeff = 0;
  int64_t old_ncoeff = 0;   // elements of the source buffer the plan may read
  std::vector<PermBox> boxes;
  std::vector<int32_t> tile0;     // per box: its first tile (prefix sum), nboxes + 1 entries
  std::vector<int32_t> tile_box;  // per tile: its box
  int64_t ntiles = 0;

  // device mirror (one fs::DeviceTables, owned), filled lazily by the
  // executor under upload_mu: boxes, then tile0, then tile_box
  void *d_tables = nullptr;
  int device = -1;
  std::mutex upload_mu;
  const PermBox *d_boxes = nullptr;
  const int32_t *d_tile0 = nullptr;
  const int32_t *d_tile_box = nullptr;
};

namespace fs {
fs_permute_plan *build_abelian_permute_plan(
    int ndim, const int64_t *shape, const int8_t *const *leg_charges,
    const uint8_t *signature, int old_nrr, int old_nblocks,
    const int8_t *old_block_irreps, const int64_t *old_block_offsets, int old_layout_f,
    int old_dense, const int32_t *axes, int new_nrr, int new_dense);
}
