"""U(1)/Z2/trivial SymmetricTensor: the reference's own assertions
(tests/test_u1_tensor.py:39-93 of frostspin) plus dense oracles. Each test runs
on the numpy emulation of the device entry points (CPU, host logic) and on the
GPU (native kernels)."""

import numpy as np
import pytest
import scipy.linalg as lg

from frostspin_amd import AsymmetricTensor, U1SymmetricTensor

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


def random_U1_tensor(row_reps, col_reps, signature, rng):
    """Dense tensor with random values on symmetric entries (reference test helper)."""
    reps = row_reps + col_reps
    irreps1D = U1SymmetricTensor.combine_representations(reps, signature)
    nnz = (irreps1D == 0).nonzero()[0]
    t0 = np.zeros(irreps1D.size)
    t0[nnz] = rng.random(nnz.size) - 0.5
    return t0.reshape(tuple(r.size for r in reps)) / lg.norm(t0)


C1 = np.array([-2, 0, 2], dtype=np.int8)
C2 = np.array([-3, -1, 1, 3], dtype=np.int8)
C3 = np.array([-2, -2, 0, 0, 0, 2, 2], dtype=np.int8)
C4 = np.array([-2, -1, 0, 1, 2], dtype=np.int8)
C5 = np.array([-3, -1, 0, 2], dtype=np.int8)
ROW = (C1, C2)
COL = (C3, C4, C5)
SIG = np.array([1, 0, 0, 1, 1], dtype=bool)


def test_reference_u1_sequence(backend):
    rng = np.random.default_rng(42)
    assert U1SymmetricTensor.symmetry() == "U1"
    assert (U1SymmetricTensor.singlet() == np.zeros((1,), dtype=np.int8)).all()
    t0 = random_U1_tensor(ROW, COL, SIG, rng)
    tu1 = U1SymmetricTensor.from_array(t0, ROW, COL, signature=SIG)
    assert (tu1.toarray() == t0).all()
    assert abs(1.0 - tu1.norm() / lg.norm(t0)) < 1e-14
    assert (tu1.transpose().toarray() == t0.transpose(2, 3, 4, 0, 1)).all()
    assert tu1.permute((0, 1), (2, 3, 4)) is tu1
    temp = tu1.permute((2, 3, 4), (0, 1))
    assert temp.nblocks == tu1.nblocks
    for b1, b2 in zip(temp.blocks, tu1.transpose().blocks, strict=True):
        assert (b1 == b2).all()
    temp = tu1.permute((3, 0, 2), (1, 4))
    assert (temp.toarray() == t0.transpose(3, 0, 2, 1, 4)).all()
    temp = tu1.permute((3,), (1, 0, 4, 2))
    assert (temp.toarray() == t0.transpose(3, 1, 0, 4, 2)).all()
    temp = tu1.dual()
    assert (temp.toarray() == t0).all()
    temp = temp.permute((3, 0, 2), (1, 4))
    assert (temp.toarray() == t0.transpose(3, 0, 2, 1, 4)).all()
    u, s, v = tu1.svd()
    assert (tu1 - (u * s.copy()) @ v).norm() < 1e-12
    assert (tu1 - u @ (s * v)).norm() < 1e-12
    s12 = s**0.5
    assert (tu1 - (u * s12) @ (s12 * v)).norm() < 1e-12
    sigma = type(tu1).from_diagonal_tensor(s)
    assert (tu1 - u @ sigma @ v).norm() < 1e-12
    tr = U1SymmetricTensor.random(ROW, COL, signature=SIG, rng=rng)
    assert abs(tr.norm() ** 2 - tr.full_contract(tr.dagger())) < 1e-12 * tr.norm() ** 2


def test_random_matches_host_generator(backend):
    """Same seed, same blocks as the reference's per-block rng.random(shape) - 0.5."""
    t = U1SymmetricTensor.random(ROW, COL, signature=SIG, rng=np.random.default_rng(7))
    rng = np.random.default_rng(7)
    irreps, shapes = U1SymmetricTensor.get_block_sizes(ROW, COL, SIG)
    assert (t.block_irreps == irreps).all()
    for b, sh in zip(t.blocks, shapes, strict=True):
        assert (b == rng.random(tuple(sh)) - 0.5).all()


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_permute_chain_and_matmul_vs_dense(backend, seed):
    rng = np.random.default_rng(seed)
    legs = [rng.integers(-2, 3, size=rng.integers(2, 5)).astype(np.int8) for _ in range(6)]
    sig = rng.integers(0, 2, size=6).astype(bool)
    t = U1SymmetricTensor.random(legs[:3], legs[3:], signature=sig, rng=rng)
    dense = t.toarray()
    for _ in range(4):
        perm = rng.permutation(6)
        nrr = int(rng.integers(1, 6))
        t2 = t.permute(tuple(perm[:nrr]), tuple(perm[nrr:]))
        assert (t2.toarray() == dense.transpose(perm)).all()
        # permute again from a transposed (column-major) tensor
        t3 = t2.transpose().permute(tuple(range(6 - nrr)), tuple(range(6 - nrr, 6)))
        assert (t3.toarray() == dense.transpose(perm).transpose(
            tuple(range(nrr, 6)) + tuple(range(nrr)))).all()
    # matmul: A (legs0..2 | legs3..5) @ B (conj legs3..5 | new legs)
    b_legs = [rng.integers(-2, 3, size=3).astype(np.int8) for _ in range(2)]
    bsig = np.concatenate([~sig[3:], rng.integers(0, 2, size=2).astype(bool)])
    b = U1SymmetricTensor.random(legs[3:], b_legs, signature=bsig, rng=rng)
    c = t @ b
    ref = t.toarray(as_matrix=True) @ b.toarray(as_matrix=True)
    assert np.abs(c.toarray(as_matrix=True) - ref).max() < 1e-13
    # matmul with dagger/transpose operands (column-major blocks)
    d = t.dagger() @ t
    refd = t.toarray(as_matrix=True).T @ t.toarray(as_matrix=True)
    assert np.abs(d.toarray(as_matrix=True) - refd).max() < 1e-13


def test_z2_int8_wraparound(backend):
    rep = np.array([-128, 0], dtype=np.int8)
    rng = np.random.default_rng(3)
    sig = np.array([False, False, True, True])
    t = U1SymmetricTensor.random((rep, rep), (rep, rep), signature=sig, rng=rng)
    assert set(t.block_irreps.tolist()) == {-128, 0}
    d = t.toarray()
    assert (t.permute((0, 2), (1, 3)).toarray() == d.transpose(0, 2, 1, 3)).all()


def test_truncated_svd_vs_scipy(backend):
    rng = np.random.default_rng(11)
    legs = [rng.integers(-1, 2, size=4).astype(np.int8) for _ in range(4)]
    t = U1SymmetricTensor.random(legs[:2], legs[2:], rng=rng)
    dense = t.toarray(as_matrix=True)
    ref = lg.svd(dense, compute_uv=False)
    u, s, v = t.truncated_svd(6)
    kept = np.sort(s.toarray())[::-1]
    assert kept.size >= 6
    assert np.abs(kept - ref[: kept.size]).max() < 1e-13 * ref[0]
    rec = (u * s) @ v
    err = np.linalg.norm(rec.toarray(as_matrix=True) - dense)
    assert abs(err - np.sqrt(np.sum(ref[kept.size :] ** 2))) < 1e-12


def test_truncated_eigh_vs_scipy(backend):
    rng = np.random.default_rng(12)
    rep = rng.integers(-2, 3, size=9).astype(np.int8)
    a = U1SymmetricTensor.random((rep,), (rep,), rng=rng)
    h = a + a.transpose().dual().dual() if False else a @ a.dagger()
    h = h - 0.3 * U1SymmetricTensor.from_array(np.eye(9), (rep,), (rep,))
    dense = h.toarray()
    ref = lg.eigvalsh(dense)
    ref = ref[np.abs(ref).argsort()[::-1]]
    s, u = h.truncated_eigh(4)
    vals = s.toarray(sort=True)
    assert np.abs(vals - ref[: vals.size]).max() < 1e-13 * np.abs(ref[0])
    rec = (u * s) @ u.dagger()
    assert np.linalg.norm(rec.toarray() - (u.toarray() * s.toarray()) @ u.toarray().T) < 1e-12


def test_asymmetric_tensor(backend):
    rng = np.random.default_rng(5)
    arr = rng.standard_normal((2, 3, 4, 5))
    reps = tuple(np.array([d]) for d in arr.shape)
    t = AsymmetricTensor.from_array(arr, reps[:2], reps[2:])
    assert (t.toarray() == arr).all()
    assert (t.permute((3, 1), (0, 2)).toarray() == arr.transpose(3, 1, 0, 2)).all()
    assert (t.transpose().toarray() == arr.transpose(2, 3, 0, 1)).all()
    m = t @ t.dagger()
    ref = arr.reshape(6, 20) @ arr.reshape(6, 20).T
    assert np.abs(m.toarray(as_matrix=True) - ref).max() < 1e-13
    u, s, v = t.svd()
    assert np.abs(s.toarray() - lg.svd(arr.reshape(6, 20), compute_uv=False)).max() < 1e-13


def test_save_load_roundtrip(backend, tmp_path):
    rng = np.random.default_rng(9)
    t = U1SymmetricTensor.random(ROW, COL, signature=SIG, rng=rng)
    f = tmp_path / "t.npz"
    t.save_to_file(f)
    t2 = U1SymmetricTensor.load_from_file(f)
    assert t2.match_representations(t)
    assert (t2.block_irreps == t.block_irreps).all()
    for b1, b2 in zip(t.blocks, t2.blocks, strict=True):
        assert (b1 == b2).all()
