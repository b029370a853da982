"""Parity against golden vectors written by the reference itself.

tests/golden/make_reference_fixtures.py imports frostspin 1.0.0 in the build
container (numba replaced by four plain-Python symbols) and saves its outputs
as npz data; the reference's own checkpoint files (SymmetricTensor,
SimpleUpdate and CTMRG npz schemas) are saved beside them. Nothing here reads
the reference at run time.

What these pin, beyond the closed-form and invariance checks elsewhere:

* SU(2) block contents for representations with external degeneracies and
  mixed signatures (in-block row order, fusion-path order, spin-reversal
  signs; lie_group_symmetric_tensor.py:856-1109) and the blocks of permutes
  and transposes (lie_group_symmetric_tensor.py:1116-1134);
* O(2) blocks with 0odd / 0even / doublets and their permutes
  (o2_symmetric_tensor.py:118-352, 853-942);
* SimpleUpdate weights, logZ and bond representations after evolve
  (simple_update.py:599-698, abstract_simple_update.py:94-147);
* CTMRG trajectories over three iterations (chi values exactly, corner
  singular values and first-neighbour rdms; ctm_renormalize.py:6-136);
* the classical Ising example at beta = 0.44, chi = 32, 200 iterations;
* loading files the reference wrote (symmetric_tensor.py:1137-1184,
  abstract_simple_update.py:149-187, abstract_ctmrg.py:168-215).

Tolerances (absolute unless stated): 1e-13 on structural transforms of
O(1) coefficients; 1e-11 on simple-update weights and logZ after 20-30
Trotter steps of QR/SVD; 1e-9 relative on CTMRG rdms and corner spectra after
three iterations (the truncated SVDs amplify rounding differences between
LAPACK and the device solvers); 1e-9 relative on the Ising energy after 200
iterations.
"""

from pathlib import Path

import numpy as np
import pytest

from frostspin_amd import O2SymmetricTensor, SU2SymmetricTensor, U1SymmetricTensor
from frostspin_amd.ctmrg import SequentialCTMRG
from frostspin_amd.ctmrg import classical_ising as ci
from frostspin_amd.simple_update import SimpleUpdate
from oracle import ising as oracle_ising

GOLDEN = Path(__file__).resolve().parent / "golden"

BACKENDS = [
    pytest.param("emulated", id="emulated"),
    pytest.param("gpu", id="gpu", marks=pytest.mark.gpu),
]

SDS = np.array(
    [[0.25, 0.0, 0.0, 0.0], [0.0, -0.25, 0.5, 0.0], [0.0, 0.5, -0.25, 0.0], [0.0, 0.0, 0.0, 0.25]]
).reshape(2, 2, 2, 2)


@pytest.fixture(params=BACKENDS)
def backend(request):
    if request.param == "emulated":
        request.getfixturevalue("emulated_device")
    else:
        request.getfixturevalue("gpu")
    return request.param


def _load(name):
    with np.load(GOLDEN / name) as fin:
        return {k: fin[k] for k in fin.files}


def _host_blocks(t):
    return list(t.blocks)


def _ref_blocks(data, prefix):
    n = data[prefix + "_block_irreps"].size
    return [data[f"{prefix}_block_{i}"] for i in range(n)]


def _ref_reps(data, prefix):
    nr = int(data[prefix + "_n_row_reps"])
    nc = int(data[prefix + "_n_col_reps"])
    return (
        [data[f"{prefix}_row_rep_{i}"] for i in range(nr)],
        [data[f"{prefix}_col_rep_{i}"] for i in range(nc)],
    )


def _assert_same_tensor(t, data, prefix, tol):
    """Representations, signature, block irreps exactly; blocks within tol."""
    rows, cols = _ref_reps(data, prefix)
    assert len(t.row_reps) == len(rows) and len(t.col_reps) == len(cols)
    for a, b in zip(t.row_reps + t.col_reps, rows + cols, strict=True):
        assert np.array_equal(np.asarray(a), b)
    assert np.array_equal(t.signature, data[prefix + "_signature"])
    assert np.array_equal(np.asarray(t.block_irreps), data[prefix + "_block_irreps"])
    for got, ref in zip(_host_blocks(t), _ref_blocks(data, prefix), strict=True):
        assert got.shape == ref.shape
        assert np.abs(got - ref).max(initial=0.0) < tol


LAYOUTS = [
    pytest.param(SU2SymmetricTensor, "ref_su2_layouts.npz", id="SU2"),
    pytest.param(O2SymmetricTensor, "ref_o2_layouts.npz", id="O2"),
]


@pytest.mark.parametrize(("cls", "fname"), LAYOUTS)
def test_block_layouts_match_reference(backend, cls, fname):
    """from_array of the reference's dense tensor gives the reference's blocks;
    the reference's saved tensor loads to the same dense array; every permute
    and transpose gives the reference's blocks."""
    data = _load(fname)
    for ic in range(int(data["n_cases"])):
        p = f"case{ic}"
        dense = data[p + "_dense"]
        rows, cols = _ref_reps(data, p + "_t")
        sig = data[p + "_t_signature"]
        t = cls.from_array(dense, rows, cols, signature=sig)
        _assert_same_tensor(t, data, p + "_t", 1e-13)
        loaded = cls.load_from_dic(data, prefix=p + "_t")
        assert np.abs(loaded.toarray() - dense).max() < 1e-13
        for ip in range(int(data[p + "_n_perms"])):
            ra = tuple(int(x) for x in data[f"{p}_perm{ip}_row_axes"])
            ca = tuple(int(x) for x in data[f"{p}_perm{ip}_col_axes"])
            tp = loaded.permute(ra, ca)
            _assert_same_tensor(tp, data, f"{p}_perm{ip}", 1e-13)
            _assert_same_tensor(tp.transpose(), data, f"{p}_perm{ip}_T", 1e-13)
            assert np.abs(tp.toarray() - dense.transpose(ra + ca)).max() < 1e-13


def _heisenberg(sym):
    if sym == "U1":
        reps = (np.array([1, -1], dtype=np.int8),) * 2
        return U1SymmetricTensor.from_array(SDS, reps, reps)
    reps = (np.array([[1], [2]]),) * 2
    return SU2SymmetricTensor.from_array(SDS, reps, reps)


def _same_rep(a, b):
    return np.array_equal(np.asarray(a), np.asarray(b))


@pytest.mark.parametrize("case", range(6))
def test_simple_update_matches_reference(backend, case):
    """Heisenberg first-neighbour simple update from beta = 0 (U(1) D = 4, 6;
    SU(2) D = 4, 7; beta 0.2 to 1): logZ, weights with multiplicities and bond
    representations equal the reference's."""
    data = _load("ref_simple_update.npz")
    p = f"case{case}"
    sym = str(data[p + "_symmetry"])
    su = SimpleUpdate.square_lattice_first_neighbor(
        _heisenberg(sym), int(data[p + "_D"]), float(data[p + "_tau"]), degen_ratio=0.99, rcutoff=1e-10
    )
    su.evolve(float(data[p + "_beta"]))
    assert abs(su.beta - float(data[p + "_beta_reached"])) < 1e-14
    assert su.Dmax == int(data[p + "_Dmax"])
    assert abs(su.logZ - float(data[p + "_logZ"])) < 1e-11
    for ib, (w, r) in enumerate(zip(su.get_weights(), su.get_bond_representations(), strict=True)):
        assert _same_rep(r, data[f"{p}_bond_rep_{ib}"])
        ref = data[f"{p}_weights_{ib}"]
        assert np.asarray(w).shape == ref.shape
        assert np.abs(np.asarray(w) - ref).max() < 1e-11


@pytest.mark.parametrize("case", range(6))
def test_load_reference_simple_update_checkpoint(backend, case):
    """A SimpleUpdate file the reference wrote loads with its weights, logZ,
    beta and bond representations, and its tensors carry those weights."""
    data = _load("ref_simple_update.npz")
    p = f"case{case}"
    su = SimpleUpdate.load_from_file(GOLDEN / f"ref_su_checkpoint_{case}.npz")
    assert su.symmetry() == str(data[p + "_symmetry"])
    assert abs(su.beta - float(data[p + "_beta_reached"])) < 1e-14
    assert su.logZ == float(data[p + "_logZ"])
    for ib, (w, r) in enumerate(zip(su.get_weights(), su.get_bond_representations(), strict=True)):
        assert _same_rep(r, data[f"{p}_bond_rep_{ib}"])
        assert np.array_equal(np.asarray(w), data[f"{p}_weights_{ib}"])
    tensors = su.get_tensors()
    assert len(tensors) == 2
    # one more step from the reference's state stays finite and keeps the reps
    su.evolve(su.tau)
    assert su.Dmax == int(data[p + "_Dmax"])


def _random_cell(seed):
    rp = np.array([1, -1], dtype=np.int8)
    ra = np.array([1, -1, 2], dtype=np.int8)
    ru = np.array([1, 1, -1, 0, 2, -2], dtype=np.int8)
    rr = np.array([2, 0, 1, -1, 0, 1, -1], dtype=np.int8)
    rd = np.array([1, 1, -1, -2], dtype=np.int8)
    rl = np.array([1, -2, 1, 0, 0], dtype=np.int8)
    sA = np.array([False, False, True, True, True, True])
    sB = np.array([False, False, False, False, False, False])
    rng = np.random.default_rng(seed)
    reps_A = (rp, ra, ru, rr, rd, rl)
    reps_B = (rp, ra, rd, rl, ru, rr)
    A = U1SymmetricTensor.random(reps_A[:2], reps_A[2:], rng=rng, signature=sA)
    B = U1SymmetricTensor.random(reps_B[:2], reps_B[2:], rng=rng, signature=sB)
    return A, B


def _compare_ctm_iteration(ctm, data, p, it, rtol):
    r21, r12 = ctm.compute_rdm_1st_neighbor_cell()
    for i, (a, b) in enumerate(zip(r21, r12, strict=True)):
        for got, key in ((a, f"{p}_it{it}_rdm2x1_{i}"), (b, f"{p}_it{it}_rdm1x2_{i}")):
            ref = data[key]
            g = got.toarray(as_matrix=True)
            assert g.shape == ref.shape
            assert np.abs(g - ref).max() <= rtol * np.abs(ref).max(), (key, np.abs(g - ref).max())
    k = 0
    for x, y in ctm.site_coords:
        for getter in (ctm.get_C1, ctm.get_C2, ctm.get_C3, ctm.get_C4):
            s = np.linalg.svd(getter(x, y).toarray(as_matrix=True), compute_uv=False)
            ref = data[f"{p}_it{it}_corner_sv_{k}"]
            assert s.shape == ref.shape
            assert np.abs(s - ref).max() <= rtol * ref[0], (p, it, k)
            k += 1


CTM_CASES = [
    pytest.param("random_u1", id="random_u1_chi20"),
    pytest.param("heis_u1_d6", id="heisenberg_u1_d6_beta0.3_chi24"),
    pytest.param("heis_su2_d7", id="heisenberg_su2_d7_beta0.3_chi24"),
    pytest.param("heis_u1_b1", id="heisenberg_u1_d4_beta1_chi24"),
    pytest.param("heis_su2_b1", id="heisenberg_su2_d4_beta1_chi24"),
]


@pytest.mark.parametrize("name", CTM_CASES)
def test_ctmrg_trajectory_matches_reference(backend, name):
    """Three SequentialCTMRG iterations from the same tensors: chi values
    exactly, corner singular values and rdms within 1e-9 relative."""
    data = _load("ref_ctmrg.npz")
    if name == "random_u1":
        tensors = _random_cell(42)
        ctm = SequentialCTMRG.from_elementary_tensors("AB\nBA", tensors, 20)
    else:
        su = SimpleUpdate.load_from_file(GOLDEN / f"ref_su_checkpoint_{int(data[name + '_su_case'])}.npz")
        ctm = SequentialCTMRG.from_elementary_tensors(
            "AB\nBA", su.get_tensors(), int(data[name + "_chi"]), degen_ratio=0.99, cutoff=1e-10
        )
    chis = data[name + "_chi_values"]
    for it in range(int(data[name + "_n_iter"])):
        ctm.iterate()
        ref = [int(c) for c in chis[it] if c > 0]
        assert [int(c) for c in ctm.chi_values] == ref, (it, ctm.chi_values, ref)
        _compare_ctm_iteration(ctm, data, name, it, 1e-9)


@pytest.mark.parametrize("name", ["random_u1", "heis_u1_d6", "heis_su2_d7", "heis_su2_b1"])
def test_load_reference_ctmrg_checkpoint(backend, name):
    """A CTMRG file the reference wrote loads into the same environment: its
    rdms and corner spectra equal the reference's at 1e-12 relative."""
    data = _load("ref_ctmrg.npz")
    ctm = SequentialCTMRG.load_from_file(GOLDEN / f"ref_ctm_{name}.npz")
    last = int(data[name + "_n_iter"]) - 1
    _compare_ctm_iteration(ctm, data, name, last, 1e-12)


@pytest.mark.parametrize("sym", ["Z2", "trivial"])
def test_ising_example_matches_reference(backend, sym):
    """The reference's classical Ising example at beta = 0.44, chi = 32, 200
    iterations: energy within 1e-9 relative, magnetization within 1e-9, and
    (Z2) the corner weights within 1e-9 and the edge tensor's transfer-matrix
    correlation length within 1e-7 relative."""
    data = _load("ref_ising.npz")
    r = ci.run(beta=float(data["beta"]), chi=int(data["chi"]), n_iter=int(data["n_iter"]),
               symmetry=sym, seed=int(data["seed"]))
    e_ref = float(data[f"energy_{sym}"])
    assert abs(r["energy"] - e_ref) < 1e-9 * abs(e_ref)
    assert abs(r["magnetization"] - float(data[f"magnetization_{sym}"])) < 1e-9
    if sym == "Z2":
        assert np.array_equal(np.asarray(r["s"].block_irreps), data["s_block_irreps"])
        for i, b in enumerate(r["s"].diagonal_blocks):
            ref = data[f"s_block_{i}"]
            assert np.asarray(b).shape == ref.shape
            assert np.abs(np.asarray(b) - ref).max() < 1e-9
        xi_ref = oracle_ising.corr_length_dense(data["T_dense"])
        assert abs(r["xi"] - xi_ref) < 1e-7 * xi_ref
