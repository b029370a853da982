"""Abelian permute on boxes larger than one LDS tile (4096 elements).

The plan holds one record per box; the kernel cuts a box into tiles along its
slowest destination digits (whole slow digits, one cut digit in chunks, the
last chunk shorter). These cases make every branch of that cut happen:
exact chunking, a last chunk of length 1 (its source entry dropped), a cut
digit merged into a faster source or destination entry, slow digits, absent
old blocks written as zeros, and column-major sources. Results are compared
bit for bit with numpy's transpose, on the numpy emulation of the device
entry points (which executes the tiles exported by the library, derived from
the boxes by the kernel's own code) and on the GPU.
"""

import numpy as np
import pytest

from frostspin_amd import AsymmetricTensor, U1SymmetricTensor, _ops

BACKENDS = [
    pytest.param("emulated", id="emulated"),
    pytest.param("gpu", id="gpu", marks=pytest.mark.gpu),
]


@pytest.fixture(params=BACKENDS)
def backend(request):
    if request.param == "emulated":
        request.getfixturevalue("emulated_device")
    else:
        request.getfixturevalue("gpu")
    return request.param


def _dense_case(shape, nrr, rng):
    arr = rng.standard_normal(shape)
    reps = tuple(np.array([d]) for d in shape)
    return arr, AsymmetricTensor.from_array(arr, reps[:nrr], reps[nrr:])


# (shape, row legs, new row axes, new column axes): one box each
DENSE_CASES = [
    ((2, 2049), 1, (1,), (0,)),  # chunk 2048, last chunk of length 1: dropped source entry
    ((3, 1500), 1, (1,), (0,)),  # chunk 1365, last chunk 135
    ((4, 2048), 1, (1,), (0,)),  # chunk 1024, exact
    ((5, 7, 300), 2, (2,), (1, 0)),  # cut digit 300, slow digits none, faster ones merged
    ((40, 30, 9), 1, (2, 0), (1,)),  # slow and cut digits, contiguous source runs
    ((9, 40, 30), 2, (1, 0), (2,)),  # destination-contiguous faster digits
    ((2, 3, 2, 1000), 2, (3,), (0, 1, 2)),  # cut digit contiguous in the source
]


@pytest.mark.parametrize("case", range(len(DENSE_CASES)))
def test_dense_boxes_larger_than_a_tile(backend, case):
    shape, nrr, rows, cols = DENSE_CASES[case]
    rng = np.random.default_rng(case)
    arr, t = _dense_case(shape, nrr, rng)
    p = t.permute(rows, cols)
    assert (p.toarray() == arr.transpose(rows + cols)).all()
    # column-major source (transposed view) through the same plan machinery
    tt = t.transpose()
    ref_t = arr.transpose(tuple(range(nrr, len(shape))) + tuple(range(nrr)))
    perm = tuple(int(x) for x in np.random.default_rng(case + 100).permutation(len(shape)))
    q = tt.permute(perm[:1], perm[1:])
    assert (q.toarray() == ref_t.transpose(perm)).all()


def test_plan_tiles_cover_every_box(backend):
    """Exported tiles: per box nchunks x slow tiles, volumes summing to the
    coefficient count, the dropped-entry case present."""
    rng = np.random.default_rng(3)
    arr, t = _dense_case((2, 2049), 1, rng)
    p = t.permute((1,), (0,))
    assert (p.toarray() == arr.T).all()
    plan = _ops.abelian_permute_plan(
        (2, 2049), [np.zeros(2, np.int8), np.zeros(2049, np.int8)], np.array([False, True]), 1,
        np.array([0], np.int8), None, False, (1, 0), 1,
    )
    tiles, nboxes = plan.tiles()
    assert nboxes == 1 and tiles.shape[0] == 2
    assert int(tiles[:, 2].sum()) == 2 * 2049
    assert sorted(tiles[:, 2].tolist()) == [2, 4096]
    assert sorted(tiles[:, 3].tolist()) == [1, 2]  # the last chunk's length-1 source entry is dropped


def test_u1_large_runs(backend):
    """U(1) legs with long runs of one charge: boxes of tens of thousands of
    elements, absent old blocks (zeros) and several permutes in a chain."""
    rng = np.random.default_rng(7)
    big = np.repeat(np.array([-1, 0, 1], dtype=np.int8), [60, 90, 50])
    small = np.array([-1, 0, 0, 1], dtype=np.int8)
    d = np.array([1, -1], dtype=np.int8)
    row, col = (big, d), (small, big)
    sig = np.array([False, True, False, True])
    t = U1SymmetricTensor.random(row, col, signature=sig, rng=rng)
    dense = t.toarray()
    for rows, cols in (((3, 0), (1, 2)), ((2,), (3, 1, 0)), ((3, 1, 0), (2,)), ((0, 3), (2, 1))):
        p = t.permute(rows, cols)
        assert (p.toarray() == dense.transpose(rows + cols)).all()
    chain = t.permute((3, 0), (1, 2)).permute((2, 0), (1, 3))
    assert (chain.toarray() == dense.transpose((3, 0, 1, 2)).transpose((2, 0, 1, 3))).all()
    # a block missing from the source: its destination boxes are zeros
    blocks = list(t.device_blocks)
    t2 = U1SymmetricTensor(row, col, blocks[1:], t.block_irreps[1:], sig)
    d2 = t2.toarray()
    p2 = t2.permute((3, 0), (1, 2))
    assert (p2.toarray() == d2.transpose((3, 0, 1, 2))).all()
