"""Truncation rule shared by truncated_svd, truncated_eigh and construct_projectors.

find_block_cuts keeps the interface of misc_tools/svd_tools.py:7-54; the greedy
loop itself runs in the native planner (fs_find_block_cuts).
"""

import numpy as np

from .._lib import find_block_cuts_native


def find_block_cuts(block_values, cut_target, *, dims=None, atol=0.0, rtol=None, degen_ratio=1.0):
    """
    Number of values to keep in each block so that about cut_target values (counted
    with dims multiplicity) survive, never splitting quasi-degenerate multiplets
    (consecutive values with ratio >= degen_ratio) and dropping values <= cutoff,
    cutoff = atol + rtol * max value, rtol defaulting to (total count) * eps.
    """
    block_values = tuple(np.asarray(t, dtype=np.float64) for t in block_values)
    n = len(block_values)
    if n == 0:
        return np.zeros((0,), dtype=np.int64)
    dims = np.ones((n,), dtype=np.int64) if dims is None else np.asarray(dims, dtype=np.int64)
    if rtol is None:
        rtol = sum(len(b) for b in block_values) * np.finfo(np.float64).eps
    heads = [b[0] for b in block_values if b.size]
    cutoff = float(atol + rtol * (max(heads) if heads else 0.0))
    degen_ratio = float(degen_ratio)
    assert dims.shape == (n,)
    assert degen_ratio <= 1.0
    return find_block_cuts_native(block_values, int(cut_target), dims, cutoff, degen_ratio)
