"""Eigensolvers: truncated_eig / eig / eigs / eigsh and the transfer-matrix
correlation length.

* The sequence of frostspin's tests/test_eigen.py: an SU(2) random matrix and
  its O(2), U(1) and trivial casts give the dense LAPACK spectrum to 1e-12
  through both the dense-block path and the device Krylov-Schur path
  (dmax_full = 10); implicit matrices (eigs with matmat), missing blocks and
  zero-block matrices; eigh / eigsh vectors satisfy M v = lambda v. Complex
  eigenvectors are not representable by the float64 device tensors: those
  calls raise NotImplementedError, checked here.
* The Krylov-Schur solver alone against numpy on operators with a prescribed
  spectrum (complex pair among the leading values, near-degenerate values).
* The transfer spectrum of one CTMRG environment is the same through the
  dense path, the Krylov path, and the environment cast to trivial symmetry;
  correlation lengths follow frostspin's tests/test_ctmrg.py checks.
"""

import numpy as np
import pytest
import scipy.linalg as lg

from frostspin_amd import SU2SymmetricTensor, U1SymmetricTensor, _ops
from frostspin_amd.ctmrg import SequentialCTMRG
from frostspin_amd.misc_tools import krylov

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


def _su2_matrix(rng):
    reps = (np.array([[2, 3, 1], [1, 3, 5]]), np.array([[2, 2, 1], [1, 3, 5]]))
    m = SU2SymmetricTensor.random(reps, reps, rng=rng)
    m /= m.norm()
    return reps, m


def _match(ref, got, tol=1e-12):
    d1 = ref[: got.size]
    return np.all((np.abs(d1 - got) < tol) | (np.abs(d1.conj() - got) < tol))


def test_truncated_eig_all_symmetries(backend):
    rng = np.random.default_rng(42)
    reps, mat_su2 = _su2_matrix(rng)
    mats = [mat_su2, mat_su2.toO2(), mat_su2.toU1(), mat_su2.totrivial()]
    default = lg.eigvals(mat_su2.toarray(as_matrix=True))
    default = default[np.abs(default).argsort()[::-1]]
    nvals, dmax_full = 20, 10
    for i, m in enumerate(mats):
        v = m.truncated_eig(nvals, dmax_full=dmax_full, rng=rng, compute_vectors=False)
        if i == 0:
            assert v.dtype == np.complex128
        v = v.toarray(sort=True)[:nvals]
        assert _match(default[:nvals], v), type(m).__name__
    # dense path only (every block below dmax_full)
    v = mat_su2.truncated_eig(nvals, dmax_full=1000, rng=rng, compute_vectors=False).toarray(sort=True)[:nvals]
    assert _match(default[:nvals], v)

    # implicit matrix M^2
    def matmat(x):
        return mat_su2 @ (mat_su2 @ x)

    vals = SU2SymmetricTensor.eigs(
        matmat, nvals, mat_su2.row_reps, mat_su2.signature[: mat_su2.n_row_reps],
        dmax_full=dmax_full, rng=rng, compute_vectors=False,
    ).toarray(sort=True)[:nvals]
    assert np.all((np.abs(default[:nvals] ** 2 - vals) < 1e-12) | (np.abs(default[:nvals].conj() ** 2 - vals) < 1e-12))


def test_eig_missing_and_zero_blocks(backend):
    rng = np.random.default_rng(42)
    reps, mat_su2 = _su2_matrix(rng)
    inds = np.array([1, 2, 4])
    mat_missing = SU2SymmetricTensor(
        reps, reps, [mat_su2.blocks[i] for i in inds], mat_su2.block_irreps[inds], mat_su2.signature
    )
    dense = mat_missing.toarray(as_matrix=True)
    ref = lg.eigvals(dense)
    ref = ref[np.abs(ref).argsort()[::-1]]
    v1 = mat_missing.truncated_eig(40, dmax_full=200, rng=rng, compute_vectors=False).toarray(sort=True)
    v2 = mat_missing.truncated_eig(4, dmax_full=2, rng=rng, compute_vectors=False).toarray(sort=True)
    assert _match(ref, v1[:40], 1e-12)
    assert _match(ref, v2[:4], 1e-12)

    def matmat(x):
        return mat_missing @ (mat_missing @ x)

    sig = mat_su2.signature[: mat_su2.n_row_reps]
    for nv, dm in ((40, 200), (4, 2)):
        vals = SU2SymmetricTensor.eigs(matmat, nv, mat_su2.row_reps, sig, dmax_full=dm, rng=rng,
                                       compute_vectors=False).toarray(sort=True)
        assert _match(ref**2, vals[:nv], 1e-12)

    mat_0b = SU2SymmetricTensor(reps, reps, [], np.array([], dtype=int), mat_su2.signature)
    assert mat_0b.norm() == 0.0
    assert mat_0b.truncated_eig(40, dmax_full=200, rng=rng, compute_vectors=False).shape == (0,)
    assert mat_0b.truncated_eig(4, dmax_full=2, rng=rng, compute_vectors=False).shape == (0,)

    def matmat0(x):
        return mat_0b @ (mat_0b @ x)

    for nv, dm in ((40, 200), (4, 2)):
        vals = SU2SymmetricTensor.eigs(matmat0, nv, mat_su2.row_reps, sig, dmax_full=dm, rng=rng,
                                       compute_vectors=False)
        assert vals.shape == (0,)


def test_eigh_eigsh_vectors(backend):
    rng = np.random.default_rng(42)
    reps, mat_su2 = _su2_matrix(rng)
    nvals, dmax_full = 20, 10
    blocks = [b + b.T for b in mat_su2.blocks]
    mat_sym = SU2SymmetricTensor(reps, reps, blocks, mat_su2.block_irreps, mat_su2.signature)
    ref = np.linalg.eigvalsh(mat_sym.toarray(as_matrix=True))
    ref = ref[np.abs(ref).argsort()[::-1]]
    vals = mat_sym.truncated_eigh(nvals, dmax_full=dmax_full, rng=rng, compute_vectors=False)
    assert vals.dtype == np.float64
    assert _match(ref, vals.toarray(sort=True)[:nvals])

    def matmat(v):
        return mat_sym @ v

    for dm in (dmax_full, 1000):
        vals = SU2SymmetricTensor.eigsh(matmat, nvals, mat_sym.row_reps, mat_sym.signature[:2], dmax_full=dm,
                                        rng=rng, compute_vectors=False)
        assert vals.dtype == np.float64
        assert _match(ref, vals.toarray(sort=True)[:nvals])

    # eigenvectors through the implicit (Krylov) path and the dense path
    for dm in (dmax_full, 1000):
        vals, vec = SU2SymmetricTensor.eigsh(matmat, nvals, mat_sym.row_reps, mat_sym.signature[:2],
                                             dmax_full=dm, rng=rng)
        assert vals.nblocks == vec.nblocks
        assert (vals.block_irreps == vec.block_irreps).all()
        assert all(vec.blocks[i].shape[1] == vals.diagonal_blocks[i].shape[0] for i in range(vals.nblocks))
        assert (vec * vals - mat_sym @ vec).norm() < 1e-12
        check = (vec.dagger() @ mat_sym @ vec).blocks
        for i, d in enumerate(vals.diagonal_blocks):
            assert np.linalg.norm(check[i] - np.diag(d)) < 1e-12

    vals, vec = mat_sym.eigh()
    assert (vec * vals - mat_sym @ vec).norm() < 1e-12
    assert (vec * vals @ vec.dagger() - mat_sym).norm() < 1e-12

    # complex eigenvectors cannot be held by float64 device tensors
    with pytest.raises(NotImplementedError):
        mat_su2.truncated_eig(nvals, dmax_full=dmax_full, rng=rng)


def test_krylov_schur_prescribed_spectrum(backend):
    rng = np.random.default_rng(7)
    d = 300
    lam = np.concatenate(([1.0, 0.9 + 0.3j, 0.9 - 0.3j, -0.95, 0.94, 0.9399999], 0.8 * rng.random(d - 6)))
    # real matrix with this spectrum: 2x2 rotation block for the complex pair
    D = np.diag(lam.real)
    D[1, 2], D[2, 1] = lam[1].imag, -lam[1].imag
    S = rng.standard_normal((d, d))
    A = S @ D @ np.linalg.inv(S)
    Ad = _ops.upload(A)

    def apply(x, y):
        _ops.grouped_gemm([(d, 1, d, 0, d, 1, 0, 1, 1, 0, 1)], Ad, x, y)

    ref = lg.eigvals(A)
    ref = ref[np.abs(ref).argsort()[::-1]]
    for k in (2, 3, 6):
        vals, _ = krylov.largest_eigenpairs(apply, d, k, _ops.upload(rng.standard_normal(d)))
        vals = np.asarray(vals)
        assert vals.size == k
        # each computed value is one of the k leading ones (pairs may come in either order)
        for v in vals:
            assert np.min(np.abs(ref[: k + 1] - v)) < 1e-10

    # symmetric operator with vectors
    Q, _ = np.linalg.qr(rng.standard_normal((d, d)))
    ev = np.concatenate(([2.0, -1.9, 1.5], rng.random(d - 3)))
    M = (Q * ev) @ Q.T
    Md = _ops.upload(M)

    def apply_sym(x, y):
        _ops.grouped_gemm([(d, 1, d, 0, d, 1, 0, 1, 1, 0, 1)], Md, x, y)

    vals, X = krylov.largest_eigenpairs(apply_sym, d, 3, _ops.upload(rng.standard_normal(d)),
                                        compute_vectors=True, symmetric=True)
    assert np.allclose(vals, [2.0, -1.9, 1.5], atol=1e-12, rtol=0)
    Xh = _ops.download(X).reshape(d, 3)
    assert np.linalg.norm(M @ Xh - Xh * vals) < 1e-10


def test_transfer_spectrum_paths_and_corr_length(backend):
    rng = np.random.default_rng(42)
    rp = np.array([1, -1], dtype=np.int8)
    rD = np.array([1, -1, 0], dtype=np.int8)
    sA = np.array([False, False, True, True, True, True])
    sB = np.zeros(6, dtype=bool)
    A = U1SymmetricTensor.random((rp, rp), (rD, rD, rD, rD), rng=rng, signature=sA)
    B = U1SymmetricTensor.random((rp, rp), (rD, rD, rD, rD), rng=rng, signature=sB)
    ctm = SequentialCTMRG.from_elementary_tensors("AB\nBA", (A, B), 12)
    for _ in range(3):
        ctm.iterate()
    chi = ctm.chi_values
    assert max(chi) > 4
    nv = 6
    for y in (0, 1):
        dense = ctm.compute_transfer_spectrum_h(y, nv, dmax_full=100000)
        kry = ctm.compute_transfer_spectrum_h(y, nv, dmax_full=4)
        assert np.all(np.isfinite(dense))
        assert _match(dense, kry, 1e-10)
    for x in (0, 1):
        dense = ctm.compute_transfer_spectrum_v(x, nv, dmax_full=100000)
        kry = ctm.compute_transfer_spectrum_v(x, nv, dmax_full=4)
        assert _match(dense, kry, 1e-10)
    xih0 = ctm.compute_corr_length_h(0)
    xih1 = ctm.compute_corr_length_h(1)
    xiv0 = ctm.compute_corr_length_v(0)
    # AB//BA: both rows (and columns) give the same correlation length
    assert xih0 > 0
    assert abs(xih1 - xih0) < 1e-10 * xih0
    assert abs(ctm.compute_corr_length_v(1) - xiv0) < 1e-10 * xiv0
    # the same environment with trivial symmetry: one big block, Krylov path
    spec_u1 = ctm.compute_transfer_spectrum_h(0, nv)
    ctm.set_symmetry("trivial")
    spec_tr = ctm.compute_transfer_spectrum_h(0, nv)
    assert _match(spec_u1, spec_tr, 1e-10)
    assert abs(ctm.compute_corr_length_h(0) - xih0) < 1e-10 * xih0
