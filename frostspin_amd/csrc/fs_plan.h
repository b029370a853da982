// Abelian permute plan: a list of box tiles built by the planner on the host,
// mirrored to device memory on first execution.
//
// A box is the set of elements whose index on every leg lies in one run of
// equal charges of that leg (a "run tuple"). Inside a box the position in the
// old block and in the new block are both affine in the legs' digits, so the
// permute of a box is a dense strided copy: a tile (a box, or a slice of a
// large box) is staged through LDS, read in source order and written in
// destination order.
#pragma once

#include <cstdint>
#include <mutex>
#include <vector>

#include "frostspin_amd.h"

// Elements of one tile (LDS staging buffer of the permute kernel).
constexpr int64_t kPermuteTileVol = 4096;
// Digits (legs with more than one state in the box) of one tile.
constexpr int kPermuteMaxDigits = FS_PERMUTE_MAX_DIGITS;

// One tile: the source is read digit by digit from the fastest source digit,
// the destination written digit by digit from the fastest destination digit.
// src_base < 0: the old block is absent, the tile is written with zeros.
// smag / dmag: ceil(2^32 / len), so that q = umulhi(f, mag) = f / len for the
// flat indices f < kPermuteTileVol of a tile.
struct PermTile {
  int64_t src_base, dst_base;
  int32_t vol, ns, nd, pad;
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
  int64_t nboxes = 0;
  std::vector<PermTile> tiles;

  // device mirror (one fs::DeviceTables, owned), filled lazily by the
  // executor under upload_mu
  void *d_tables = nullptr;
  int device = -1;
  std::mutex upload_mu;
  const PermTile *d_tiles = nullptr;
};

namespace fs {
fs_permute_plan *build_abelian_permute_plan(
    int ndim, const int64_t *shape, const int8_t *const *leg_charges,
    const uint8_t *signature, int old_nrr, int old_nblocks,
    const int8_t *old_block_irreps, const int64_t *old_block_offsets, int old_layout_f,
    int old_dense, const int32_t *axes, int new_nrr, int new_dense);
}
