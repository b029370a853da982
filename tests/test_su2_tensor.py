"""SU(2) SymmetricTensor: the reference's own assertions (tests/test_su2_tensor.py
of frostspin) and parity of every structural transform (from_array, toarray,
permute, transpose) with the SU(2) oracle. Each test runs on the numpy emulation
of the device entry points (host logic and plan tables) and on the GPU (native
structural-transform kernel). Tolerance: 1e-13 absolute on blocks of O(1)
coefficients (the transforms are sums of a few tens of products)."""

import numpy as np
import pytest
import scipy.linalg as lg

from frostspin_amd import SU2SymmetricTensor, U1SymmetricTensor
from oracle import su2 as osu2

BACKENDS = [
    pytest.param("emulated", id="emulated"),
    pytest.param("gpu", id="gpu", marks=pytest.mark.gpu),
]
TOL = 1e-13


@pytest.fixture(params=BACKENDS)
def backend(request):
    if request.param == "emulated":
        request.getfixturevalue("emulated_device")
    else:
        request.getfixturevalue("gpu")
    return request.param


def host_blocks(t):
    return list(t.blocks)


def assert_matches(t, o, tol=TOL):
    assert list(t.block_irreps) == list(o.block_irreps)
    tb = host_blocks(t)
    assert len(tb) == len(o.blocks)
    for a, b in zip(tb, o.blocks, strict=True):
        assert a.shape == b.shape
        assert np.abs(a - b).max(initial=0.0) < tol


SDS_22 = np.array(
    [[0.25, 0.0, 0.0, 0.0], [0.0, -0.25, 0.5, 0.0], [0.0, 0.5, -0.25, 0.0], [0.0, 0.0, 0.0, 0.25]]
)
SDS_22B = np.array(
    [[-0.25, 0.0, 0.0, -0.5], [0.0, 0.25, 0.0, 0.0], [0.0, 0.0, 0.25, 0.0], [-0.5, 0.0, 0.0, -0.25]]
)
R = np.array([[1], [2]])


def test_reference_su2_sequence(backend):
    """test_su2_tensor.py:17-109 of the reference."""
    assert SU2SymmetricTensor.symmetry() == "SU2"
    assert (SU2SymmetricTensor.singlet() == np.array([[1], [1]])).all()
    t = SDS_22.reshape(2, 2, 2, 2)
    st = SU2SymmetricTensor.from_array(t, (R, R), (R, R))
    assert st.nblocks == 2
    assert (st.block_irreps == np.array([1, 3])).all()
    b = host_blocks(st)
    assert b[0].shape == (1, 1) and b[1].shape == (1, 1)
    assert abs(b[0][0, 0] + 0.75) < 1e-14
    assert abs(b[1][0, 0] - 0.25) < 1e-14
    assert abs(lg.norm(SDS_22) - st.norm()) < 1e-14
    assert lg.norm(st.toarray() - t) < 1e-14
    assert lg.norm(st.dual().toarray() - t) < 1e-14
    assert abs(st.trace()) < 1e-14

    stp = st.permute((0, 1, 2), (3,))
    stp2 = SU2SymmetricTensor.from_array(t, (R, R, R), (R,), signature=[0, 0, 1, 1])
    assert (stp - stp2).norm() < 1e-14
    assert (st - stp.permute((0, 1), (2, 3))).norm() < 1e-14
    stp = st.permute((3, 0, 2), (1,))
    stp2 = SU2SymmetricTensor.from_array(t.transpose(3, 0, 2, 1), (R, R, R), (R,), signature=[1, 0, 1, 0])
    assert (stp - stp2).norm() < 1e-14
    assert (st - stp.permute((1, 3), (2, 0))).norm() < 1e-14

    tb = SDS_22B.reshape(2, 2, 2, 2)
    stb = SU2SymmetricTensor.from_array(tb, (R, R), (R, R), signature=[0, 1, 1, 0])
    bb = host_blocks(stb)
    assert abs(bb[0][0, 0] + 0.75) < 1e-14 and abs(bb[1][0, 0] - 0.25) < 1e-14
    assert lg.norm(stb.toarray() - tb) < 1e-14
    stbp = stb.permute((0, 1, 2), (3,))
    stbp2 = SU2SymmetricTensor.from_array(tb, (R, R, R), (R,), signature=[0, 1, 1, 0])
    assert (stbp - stbp2).norm() < 1e-14
    stbp = stb.permute((3, 0, 2), (1,))
    stbp2 = SU2SymmetricTensor.from_array(tb.transpose(3, 0, 2, 1), (R, R, R), (R,), signature=[0, 0, 1, 1])
    assert (stbp - stbp2).norm() < 1e-14
    assert (stb - stbp.permute((1, 3), (2, 0))).norm() < 1e-14

    # external degeneracies
    r1 = np.array([[1], [1]])
    r12 = np.array([[1, 2, 1], [1, 3, 5]])
    arr = np.eye(12).reshape(1, 12, 12, 1)
    st3 = SU2SymmetricTensor.from_array(arr, (r1,), (r12, r12, r1), signature=[0, 0, 1, 1])
    assert lg.norm(st3.toarray() - arr) < 1e-14
    st3p = st3.permute((0, 1), (2, 3))
    assert lg.norm(st3p.toarray() - arr) < 1e-14

    # cast
    assert lg.norm(st.toU1().toarray() - t) < 1e-14
    assert isinstance(st.toU1(), U1SymmetricTensor)

    # scalar operations
    assert lg.norm((st / 2).toarray() - t / 2) < 1e-14
    assert lg.norm((st * 2).toarray() - t * 2) < 1e-14
    assert lg.norm((2 * st).toarray() - t * 2) < 1e-14
    st /= 2
    assert lg.norm(st.toarray() - t / 2) < 1e-14
    st *= 2
    assert lg.norm(st.toarray() - t) < 1e-14


def test_reference_missing_blocks(backend):
    """test_su2_tensor.py:111-123: a product with a single block among many allowed."""
    r3 = np.array([[1], [3]])
    t = SU2SymmetricTensor.random((r3, r3, r3), (r3,), rng=np.random.default_rng(0))
    t = t @ t.dagger()
    assert t.nblocks == 1
    tp = t.permute((3, 2, 1), (0, 5, 4))
    tp2 = tp.permute((3, 2, 1), (0, 5, 4))
    assert (tp2 - t).norm() < 1e-14
    t2 = SU2SymmetricTensor.from_array(t.toarray(), t.row_reps, t.col_reps, t.signature)
    assert (t - t2).norm() < 1e-14


D8 = np.array([[1, 2, 1], [1, 2, 3]])  # spin 0 + 2 x spin 1/2 + spin 1 (D = 8)
CHI = np.array([[2, 3, 2, 1], [1, 2, 3, 4]])
CASES = [
    ((D8, CHI), (D8, R), [0, 1, 1, 0], [((3, 0), (2, 1)), ((1,), (0, 2, 3)), ((2, 3), (0, 1))]),
    ((R, R, D8), (CHI, D8), [1, 0, 0, 1, 0], [((0, 4, 2), (3, 1)), ((3, 2, 1, 0), (4,)), ((4, 0), (1, 2, 3))]),
    ((CHI,), (D8, R, D8), [0, 1, 1, 0], [((1, 2), (0, 3)), ((3, 0, 1), (2,))]),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_structural_transforms_match_oracle(backend, case):
    rows, cols, sig, perms = CASES[case]
    sig = np.array(sig, dtype=bool)
    o = osu2.SU2Tensor.random(rows, cols, sig, np.random.default_rng(100 + case))
    dense = o.toarray()
    t = SU2SymmetricTensor.from_array(dense, rows, cols, sig)
    assert_matches(t, o)
    assert np.abs(t.toarray() - dense).max() < TOL
    for row_axes, col_axes in perms:
        assert_matches(t.permute(row_axes, col_axes), o.permute(row_axes, col_axes))
    assert_matches(t.transpose(), o.transpose())
    # Fortran-order (dagger) input and permute of the result
    row_axes, col_axes = perms[0]
    assert_matches(t.dagger().permute(col_axes, row_axes), o.dagger().permute(col_axes, row_axes))
    assert abs(t.norm() - o.norm()) < TOL


def test_linear_algebra_matches_oracle(backend):
    rng = np.random.default_rng(5)
    sa = np.array([0, 0, 1, 1], dtype=bool)
    A = osu2.SU2Tensor.random((D8, CHI), (D8, R), sa, rng)
    B = osu2.SU2Tensor.random((D8, R), (CHI,), np.array([0, 0, 1], dtype=bool), rng)
    tA = SU2SymmetricTensor.from_array(A.toarray(), A.row_reps, A.col_reps, A.signature)
    tB = SU2SymmetricTensor.from_array(B.toarray(), B.row_reps, B.col_reps, B.signature)
    assert_matches(tA @ tB, A @ B, tol=1e-12)
    assert abs((tA @ tA.dagger()).trace() - (A @ A.dagger()).trace()) < 1e-11
    u, s, v = tA.truncated_svd(30)
    ou, os_, oirr, ov = A.truncated_svd(30)
    assert list(s.block_irreps) == list(oirr)
    for a, b in zip(s.diagonal_blocks, os_, strict=True):
        assert np.abs(a - b).max() < 1e-12
    u, s, v = tA.truncated_svd(10**6)
    assert (u * s @ v - tA).norm() < 1e-12 * tA.norm()
