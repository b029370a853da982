"""Allocation-free grouped GEMM (fs_grouped_gemm_f64_noalloc) against the
allocating entry point (fs_grouped_gemm_f64) on the same inputs, called through
the C-ABI.

Both entry points launch the same kernels in the same order (split-K partials
summed in split order), so the outputs must be bit-identical; both are also
checked against a host FP64 product (1e-12 relative, the GEMM's own test
tolerance). Two calls in a row, the second with smaller inputs; every call
takes the split-K path (few output tiles over a long K), so the static scratch
array is written by the first call and reused by the second; the first call
also mixes uneven groups and a skinny-row group."""

import ctypes

import numpy as np
import pytest

from frostspin_amd import _lib, _ops


def _layout(shapes):
    """Groups (m, n, k) packed back to back: A_i (m x k), B_i (k x n), C_i (m x n)
    row-major; returns the fs_gemm_group table and the buffer sizes."""
    rows, a_off, b_off, c_off = [], 0, 0, 0
    for m, n, k in shapes:
        rows.append((m, n, k, a_off, k, 1, b_off, n, 1, c_off, n))
        a_off += m * k
        b_off += k * n
        c_off += m * n
    return np.array(rows, dtype=np.int64), a_off, b_off, c_off


def _call(fn, table, A, B, C):
    arr = table.ctypes.data_as(ctypes.POINTER(_lib.GemmGroup))
    status = fn(len(table), arr, _ops._ptr(A), _ops._ptr(B), _ops._ptr(C), 1.0, 0.0, _ops._stream())
    assert status == 0, _lib.lib.fs_last_error()


CASES = [
    # uneven tiled groups (split over K in 11) plus one skinny-row group
    [(200, 150, 3000), (64, 500, 1024), (17, 33, 700), (8, 1024, 64)],
    # one small product with a long K (split over K in 8), smaller than the first call
    [(64, 64, 2048)],
]


@pytest.mark.gpu
def test_noalloc_gemm_matches_allocating_entry_point(gpu):
    import torch

    rng = np.random.default_rng(7)
    lib = _lib.lib
    for shapes in CASES:
        table, na, nb, nc = _layout(shapes)
        a_host = rng.standard_normal(na)
        b_host = rng.standard_normal(nb)
        A, B = _ops.upload(a_host), _ops.upload(b_host)
        C_ref = _ops.zeros(nc)
        C_new = torch.full((nc,), np.nan, dtype=torch.float64, device=gpu)
        _call(lib.fs_grouped_gemm_f64, table, A, B, C_ref)
        _call(lib.fs_grouped_gemm_f64_noalloc, table, A, B, C_new)
        torch.cuda.synchronize()
        ref = C_ref.cpu().numpy()
        got = C_new.cpu().numpy()
        assert np.array_equal(got, ref)
        for m, n, k, ao, _, _, bo, _, _, co, _ in table:
            host = a_host[ao : ao + m * k].reshape(m, k) @ b_host[bo : bo + k * n].reshape(k, n)
            assert np.abs(got[co : co + m * n].reshape(m, n) - host).max() <= 1e-12 * np.abs(host).max()
