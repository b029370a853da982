"""O(2) SymmetricTensor: the reference's own assertions (tests/test_o2_tensor.py
and the O(2)/SU(2)/U(1) RVB cross-checks of tests/test_RVB.py of frostspin) and
parity of from_array / toarray / toU1 / permute / transpose with the O(2)
oracle. Emulated (CPU) and GPU backends. Tolerance 1e-14 absolute on O(1)
coefficients (splits and merges are +-1/sqrt(2) combinations)."""

import numpy as np
import pytest
import scipy.linalg as lg

from frostspin_amd import O2SymmetricTensor, SU2SymmetricTensor, U1SymmetricTensor
from frostspin_amd.ctmrg import SequentialCTMRG
from oracle import o2 as oo2

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


R1 = np.array([[1, 1, 1, 1], [-1, 0, 1, 2]])
R2 = np.array([[2, 1, 1, 1], [-1, 0, 1, 2]])
R3 = np.array([[1, 2, 1, 1], [-1, 0, 1, 2]])
R4 = np.array([[1, 1, 1, 1], [-1, 0, 2, 3]])
R5 = np.array([[2, 1, 1, 1], [-1, 0, 2, 3]])


def hb(t):
    return list(t.blocks)


def check_o2(to2, tu1):
    """test_o2_tensor.py:18-32 of the reference."""
    assert lg.norm(tu1.toarray() - to2.toarray()) < 1e-15
    temp = O2SymmetricTensor.from_U1(tu1, to2.row_reps, to2.col_reps)
    assert (temp - to2).norm() < 1e-15
    row_axes, col_axes = (2, 1, 4), (0, 3)
    tpo2 = to2.permute(row_axes, col_axes)
    tpu1 = tu1.permute(row_axes, col_axes)
    assert (tpo2.toU1() - tpu1).norm() < 1e-15
    assert lg.norm(tpo2.toarray() - tpu1.toarray()) < 1e-15
    tpo2 = to2.transpose()
    tpu1 = tu1.transpose()
    assert (tpo2.toU1() - tpu1).norm() < 1e-15
    assert lg.norm(tpo2.toarray() - tpu1.toarray()) < 1e-15


def test_reference_o2_sequence(backend):
    rng = np.random.default_rng(42)
    assert O2SymmetricTensor.symmetry() == "O2"
    assert (O2SymmetricTensor.singlet() == np.array([[1], [0]])).all()
    to2 = O2SymmetricTensor.random((R1, R2, R3), (R4, R5), rng=rng)
    to2 /= to2.norm()
    tu1 = to2.toU1()
    check_o2(to2, tu1)
    # pathological cases: missing 0odd / 0even / both / everything but them
    t2 = to2.copy()
    t2.device_blocks[0][:] = 0
    t2u1 = t2.toU1()
    t2 = O2SymmetricTensor(t2.row_reps, t2.col_reps, t2.blocks[1:], t2.block_irreps[1:], t2.signature)
    check_o2(t2, t2u1)
    t2 = to2.copy()
    t2.device_blocks[1][:] = 0
    t2u1 = t2.toU1()
    t2 = O2SymmetricTensor(t2.row_reps, t2.col_reps, t2.blocks[:1] + t2.blocks[2:],
                           (-1, *t2.block_irreps[2:]), to2.signature)
    check_o2(t2, t2u1)
    t2 = to2.copy()
    t2.device_blocks[0][:] = 0
    t2.device_blocks[1][:] = 0
    t2u1 = t2.toU1()
    t2 = O2SymmetricTensor(t2.row_reps, t2.col_reps, t2.blocks[2:], t2.block_irreps[2:], t2.signature)
    check_o2(t2, t2u1)
    t2 = O2SymmetricTensor(to2.row_reps, to2.col_reps, to2.blocks[:2], to2.block_irreps[:2], to2.signature)
    i0 = int(np.searchsorted(tu1.block_irreps, 0))
    t2u1 = U1SymmetricTensor(tu1.row_reps, tu1.col_reps, (tu1.blocks[i0],), (0,), tu1.signature)
    check_o2(t2, t2u1)
    # only doublets / only odd / only even fixed points in the columns
    for r6, r7 in (
        (np.array([[1, 1, 1], [2, 3, 4]]), np.array([[2, 1], [2, 3]])),
        (np.array([[4], [-1]]), np.array([[5], [-1]])),
        (np.array([[4], [0]]), np.array([[5], [0]])),
    ):
        to2 = O2SymmetricTensor.random((R1, R2, r6), (R3, r7), rng=rng)
        to2 /= to2.norm()
        tu1 = to2.toU1()
        to2 = to2.permute((0, 1, 3), (2, 4))
        tu1 = tu1.permute((0, 1, 3), (2, 4))
        check_o2(to2, tu1)


def test_o2_matches_oracle(backend):
    sig = np.array([0, 0, 0, 1, 1], dtype=bool)
    o = oo2.O2Tensor.random((R1, R2, R3), (R4, R5), sig, np.random.default_rng(3))
    d = o.toarray()
    t = O2SymmetricTensor.from_array(d, (R1, R2, R3), (R4, R5), sig)
    assert list(t.block_irreps) == list(o.block_irreps)
    for a, b in zip(hb(t), o.blocks, strict=True):
        assert np.abs(a - b).max() < 1e-14
    assert np.abs(t.toarray() - d).max() < 1e-14
    tu, ou = t.toU1(), o.to_u1()
    assert list(tu.block_irreps) == list(ou.block_irreps)
    for a, b in zip(hb(tu), ou.blocks, strict=True):
        assert np.abs(a - b).max() < 1e-14
    for axes in (((2, 1, 4), (0, 3)), ((4,), (3, 2, 1, 0)), ((0, 1, 3), (2, 4))):
        tp, op = t.permute(*axes), o.permute(*axes)
        assert list(tp.block_irreps) == list(op.block_irreps)
        for a, b in zip(hb(tp), op.blocks, strict=True):
            assert np.abs(a - b).max() < 1e-14


def _rvb():
    d, D = 2, 3
    t = np.zeros((d, 1, D, D, D, D))
    for i in range(d):
        t[i, 0, 0, 0, 0, i + 1] = 1.0
        t[i, 0, 0, 0, i + 1, 0] = 1.0
        t[i, 0, 0, i + 1, 0, 0] = 1.0
        t[i, 0, i + 1, 0, 0, 0] = 1.0
    rr_u1 = (np.array([1, -1], dtype=np.int8), np.array([0], dtype=np.int8))
    rc_u1 = (np.array([0, 1, -1], dtype=np.int8),) * 4
    rr_o2 = (np.array([[1], [1]]), np.array([[1], [0]]))
    rc_o2 = (np.array([[1, 1], [0, 1]]),) * 4
    rr_su2 = (np.array([[1], [2]]), np.array([[1], [1]]))
    rc_su2 = (np.array([[1, 1], [1, 2]]),) * 4
    return t, (rr_u1, rc_u1), (rr_o2, rc_o2), (rr_su2, rc_su2)


def test_reference_rvb_cross_symmetry(backend):
    """test_RVB.py:37-91 of the reference."""
    tol = 5e-14
    t, (ru, cu), (ro, co), (rs, cs) = _rvb()
    for sig in ([0, 0, 1, 1, 1, 1], [1, 1, 0, 0, 0, 0]):
        tu = U1SymmetricTensor.from_array(t, ru, cu, signature=sig)
        to = O2SymmetricTensor.from_array(t, ro, co, signature=sig)
        ts = SU2SymmetricTensor.from_array(t, rs, cs, signature=sig)
        assert lg.norm(to.toarray() - t) < tol
        assert lg.norm(ts.toarray() - t) < tol
        assert (to.toU1() - tu).norm() < tol
        assert (ts.toU1() - tu).norm() < tol
    tu = U1SymmetricTensor.from_array(t, ru, cu, signature=[0, 0, 1, 1, 1, 1])
    to = O2SymmetricTensor.from_array(t, ro, co, signature=[0, 0, 1, 1, 1, 1])
    ts = SU2SymmetricTensor.from_array(t, rs, cs, signature=[0, 0, 1, 1, 1, 1])
    a0 = np.tensordot(t, t, ((0, 1), (0, 1)))
    au, ao, as_ = tu.dagger() @ tu, to.dagger() @ to, ts.dagger() @ ts
    assert lg.norm(au.toarray() - a0) < tol
    assert lg.norm(ao.toarray() - a0) < tol
    assert lg.norm(as_.toarray() - a0) < tol
    assert (ao.toU1() - au).norm() < tol
    assert (as_.toU1() - au).norm() < tol
    a1 = a0.transpose(0, 4, 1, 5, 2, 6, 3, 7)
    perm = ((0, 4, 1, 5), (2, 6, 3, 7))
    a1u, a1o, a1s = au.permute(*perm), ao.permute(*perm), as_.permute(*perm)
    assert lg.norm(a1u.toarray() - a1) < tol
    assert lg.norm(a1o.toarray() - a1) < tol
    assert lg.norm(a1s.toarray() - a1) < tol
    assert (a1o.toU1() - a1u).norm() < tol
    assert (a1s.toU1() - a1u).norm() < tol
    # SU(2) -> O(2) cast keeps the dense tensor up to leg basis conventions
    assert abs(ts.toO2().norm() - ts.norm()) < tol


def _check_rdm(c1, c2, tol):
    r1 = c1.compute_rdm2x1(0, 0).toarray(as_matrix=True)
    r2 = c2.compute_rdm2x1(0, 0).toarray(as_matrix=True)
    assert lg.norm(r1 - r1.T) < tol
    assert lg.norm(r2 - r2.T) < tol
    assert lg.norm(lg.eigvalsh(r1) - lg.eigvalsh(r2)) < tol


def test_reference_rvb_ctmrg(backend):
    """test_RVB.py:99-116: CTMRG on the RVB state, U(1) against O(2) (cast from
    SU(2)) and against SU(2): same reduced density matrix spectrum for 4 iterations."""
    tol = 5e-14
    t, (ru, cu), _, (rs, cs) = _rvb()
    tA_u = U1SymmetricTensor.from_array(t, ru, cu, signature=[0, 0, 1, 1, 1, 1])
    tB_u = U1SymmetricTensor.from_array(t, ru, cu, signature=[1, 1, 0, 0, 0, 0])
    tA_s = SU2SymmetricTensor.from_array(t, rs, cs, signature=[0, 0, 1, 1, 1, 1])
    tB_s = SU2SymmetricTensor.from_array(t, rs, cs, signature=[1, 1, 0, 0, 0, 0])
    for other in ("O2", "SU2"):
        cu = SequentialCTMRG.from_elementary_tensors("AB\nBA", (tA_u, tB_u), 20, degen_ratio=0.99)
        co = SequentialCTMRG.from_elementary_tensors("AB\nBA", (tA_s, tB_s), 20, degen_ratio=0.99)
        if other == "O2":
            co.set_symmetry("O2")
        _check_rdm(cu, co, tol)
        for _ in range(4):
            cu.iterate()
            co.iterate()
            _check_rdm(cu, co, tol)
