// Permute plan: host tables built by the planner, mirrored to device memory
// on first execution.
#pragma once

#include <cstdint>
#include <mutex>
#include <vector>

#include "frostspin_amd.h"

// Elements handled by one workgroup of the permute gather kernel.
constexpr int64_t kPermuteTile = 2048;

struct fs_permute_plan {
  // old side (per dense row / column of the old matrix view)
  std::vector<int32_t> old_row_block, old_row_pos, old_col_block, old_col_pos;
  // old side (per old block)
  std::vector<int64_t> old_off, old_rs, old_cs;
  int64_t old_ncoeff = 0;
  // new side (per new block)
  std::vector<int8_t> new_irreps;
  std::vector<int64_t> new_shapes;  // 2 per block
  std::vector<int64_t> new_off, new_rowstart, new_colstart;
  // new side (per in-block row / column): old (row, col) index contributions
  std::vector<int32_t> row_a, col_b;
  int64_t new_ncoeff = 0;
  int64_t n_old_rows = 0, n_old_cols = 0, n_new_rows_total = 0, n_new_cols_total = 0;
  std::vector<int64_t> tile_prefix;

  // device mirror (one fs::DeviceTables, owned), filled lazily by the
  // executor under upload_mu
  void *d_tables = nullptr;
  int device = -1;
  std::mutex upload_mu;
  const int32_t *d_old_row_block = nullptr, *d_old_row_pos = nullptr;
  const int32_t *d_old_col_block = nullptr, *d_old_col_pos = nullptr;
  const int32_t *d_row_a = nullptr, *d_col_b = nullptr;
  const int64_t *d_old_off = nullptr, *d_old_rs = nullptr, *d_old_cs = nullptr;
  const int64_t *d_new_off = nullptr, *d_new_rowstart = nullptr;
  const int64_t *d_new_colstart = nullptr, *d_new_shapes = nullptr;
  const int64_t *d_tile_prefix = nullptr;

  void finalize_tiles();
};

namespace fs {
fs_permute_plan *build_abelian_permute_plan(
    int ndim, const int64_t *shape, const int8_t *const *leg_charges,
    const uint8_t *signature, int old_nrr, int old_nblocks,
    const int8_t *old_block_irreps, const int64_t *old_block_offsets, int old_layout_f,
    int old_dense, const int32_t *axes, int new_nrr, int new_dense);
}
