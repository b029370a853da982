"""SimultaneousCTMRG, SymmetrizedCTMRG and SymmetrizedSimpleUpdate: the
sequence of frostspin's tests/test_symmetrized_ctm.py with its tolerances.

* On an AB/BA cell with B the rotated conjugate of A, the symmetrized and
  simultaneous schedules hold the same environment before any iteration
  (1e-14) and the same gauge-invariant quantities after (corner singular
  values, rdms, PEPS norm).
* SymmetrizedSimpleUpdate on the rotated Heisenberg bond (one stored tensor)
  gives the weights and bond representations of the two-tensor SimpleUpdate;
  the three CTMRG schedules then agree on rdm spectra and on the diagonal
  (next-nearest-neighbour) rdms.
* save_to_file / load_from_file round trips of both simple-update classes.

The emulated run is a reduced sequence (fewer iterations and smaller chi) so
that the CPU suite stays short; the GPU run is the full reference sequence.
"""

import numpy as np
import pytest

from frostspin_amd import SU2SymmetricTensor
from frostspin_amd.ctmrg import SequentialCTMRG, SimultaneousCTMRG, SymmetrizedCTMRG
from frostspin_amd.simple_update import SimpleUpdate, SymmetrizedSimpleUpdate

ASSERT_TOL = 1e-14

# chi of the random-state run and its iterations after the first; simple-update
# D, CTMRG chi and iterations, successive evolve() lengths
FULL = {"random_chi": 50, "extra_iters": 4, "D": 7, "chi": 49, "ctm_iters": 10, "betas": (0.5, 1.0)}
REDUCED = {"random_chi": 20, "extra_iters": 1, "D": 4, "chi": 16, "ctm_iters": 8, "betas": (0.1,)}

BACKENDS = [
    pytest.param("emulated", id="emulated"),
    pytest.param("gpu", id="gpu", marks=pytest.mark.gpu),
]


@pytest.fixture(params=BACKENDS)
def backend(request):
    if request.param == "emulated":
        request.getfixturevalue("emulated_device")
        return REDUCED
    request.getfixturevalue("gpu")
    return FULL


def _svdvals(st):
    return st.svd()[1]


def test_symmetrized_matches_simultaneous_random_state(backend):
    rng = np.random.default_rng(42)
    r2 = np.array([[1], [2]])
    rDu = np.array([[2, 3], [1, 3]])
    rDr = np.array([[3, 3], [1, 3]])
    rDd = np.array([[1, 3], [1, 3]])
    rDl = np.array([[4, 3], [1, 3]])
    A = SU2SymmetricTensor.random((r2, r2), (rDu, rDr, rDd, rDl), rng=rng)
    A /= A.norm()
    B = A.dagger().permute((4, 5), (2, 3, 0, 1))
    chi = backend["random_chi"]
    ctm = SimultaneousCTMRG.from_elementary_tensors("AB\nBA", [A, B], chi)
    ctm2 = SymmetrizedCTMRG.from_elementary_tensors("AB\nBA", [A, B], chi)
    getters = ("get_A", "get_T1", "get_T2", "get_T3", "get_T4", "get_C1", "get_C2", "get_C3", "get_C4")
    corners = ("construct_enlarged_dl", "construct_enlarged_ul", "construct_enlarged_ur", "construct_enlarged_dr")
    for coords in [(0, 0), (0, 1)]:
        for g in getters + corners:
            assert (getattr(ctm, g)(*coords) - getattr(ctm2, g)(*coords)).norm() < ASSERT_TOL, (g, coords)
    assert np.isclose(ctm.compute_PEPS_norm_log(), ctm2.compute_PEPS_norm_log())

    ctm.iterate()
    ctm2.iterate()
    # SVD gauge freedom: compare gauge-invariant quantities only
    for coords in [(0, 0), (0, 1)]:
        for g in corners:
            s1 = _svdvals(getattr(ctm, g)(*coords))
            s2 = _svdvals(getattr(ctm2, g)(*coords))
            assert (s1 - s2).norm() < 100 * ASSERT_TOL * s1.norm(), (g, coords)
        assert (ctm.compute_rdm1x2(*coords) - ctm2.compute_rdm1x2(*coords)).norm() < ASSERT_TOL
        assert (ctm.compute_rdm2x1(*coords) - ctm2.compute_rdm2x1(*coords)).norm() < ASSERT_TOL
    assert np.isclose(ctm.compute_PEPS_norm_log(), ctm2.compute_PEPS_norm_log())

    for _ in range(backend["extra_iters"]):
        ctm.iterate()
        ctm2.iterate()
        for coords in [(0, 0), (0, 1)]:
            assert (ctm.compute_rdm1x2(*coords) - ctm2.compute_rdm1x2(*coords)).norm() < ASSERT_TOL
            assert (ctm.compute_rdm2x1(*coords) - ctm2.compute_rdm2x1(*coords)).norm() < ASSERT_TOL
        assert np.isclose(ctm.compute_PEPS_norm_log(), ctm2.compute_PEPS_norm_log())


def _hamiltonians():
    sds22 = np.array([[0.25, 0, 0, 0], [0, -0.25, 0.5, 0], [0, 0.5, -0.25, 0], [0, 0, 0, 0.25]])
    # bond operator after rotating the second site's basis (sublattice rotation)
    sds22b = np.array([[-0.25, 0, 0, -0.5], [0, 0.25, 0, 0], [0, 0, 0.25, 0], [-0.5, 0, 0, -0.25]])
    r2 = np.array([[1], [2]])
    h22 = SU2SymmetricTensor.from_array(sds22.reshape(2, 2, 2, 2), (r2, r2), (r2, r2))
    h22b = SU2SymmetricTensor.from_array(sds22b.reshape(2, 2, 2, 2), (r2, r2), (r2, r2), [0, 1, 1, 0])
    return h22, h22b


def _compare_weights(su0, su):
    reps0 = su0.get_bond_representations()
    reps = su.get_bond_representations()
    weights0 = su0.get_weights()
    weights = su.get_weights()
    assert len(reps) == 4
    assert len(weights) == 4
    for i in range(4):
        assert reps0[i].shape == reps[i].shape
        assert (reps0[i] == reps[i]).all()
        assert np.allclose(weights0[i], weights[i])


def _compare_schedules(su0, su, chi, iters):
    ctm0 = SequentialCTMRG.from_elementary_tensors("AB\nBA", su0.get_tensors(), chi)
    ctm1 = SimultaneousCTMRG.from_elementary_tensors("AB\nBA", su.get_tensors(), chi)
    ctm2 = SymmetrizedCTMRG.from_elementary_tensors("AB\nBA", su.get_tensors(), chi)
    for _ in range(iters):
        ctm0.iterate()
        ctm1.iterate()
        ctm2.iterate()
    assert np.isclose(ctm0.compute_PEPS_norm_log(), ctm1.compute_PEPS_norm_log())
    assert np.isclose(ctm1.compute_PEPS_norm_log(), ctm2.compute_PEPS_norm_log())
    for coords in [(0, 0), (0, 1)]:
        for m in ("compute_rdm1x2", "compute_rdm2x1"):
            sp0 = getattr(ctm0, m)(*coords).eigh(compute_vectors=False)
            sp1 = getattr(ctm1, m)(*coords).eigh(compute_vectors=False)
            sp2 = getattr(ctm2, m)(*coords).eigh(compute_vectors=False)
            assert (sp0 - sp1).norm() < 1e-7  # different simple updates: low precision
            assert (sp1 - sp2).norm() < 10 * ASSERT_TOL

    rdm0_dr, rdm0_ur = ctm0.compute_rdm_2nd_neighbor_cell()
    rdm1_dr, rdm1_ur = ctm1.compute_rdm_2nd_neighbor_cell()
    rdm2_dr, rdm2_ur = ctm2.compute_rdm_2nd_neighbor_cell()
    assert (rdm0_dr[0] - rdm1_dr[0]).norm() < 1e-7
    assert (rdm0_dr[1] - rdm1_dr[1].transpose()).norm() < 1e-7
    assert (rdm1_dr[0] - rdm2_dr[0]).norm() < ASSERT_TOL
    assert (rdm1_dr[1] - rdm2_dr[1]).norm() < ASSERT_TOL
    assert (rdm0_ur[0] - rdm1_ur[0].transpose()).norm() < 1e-7
    assert (rdm0_ur[1] - rdm1_ur[1]).norm() < 1e-7
    assert (rdm1_ur[0] - rdm2_ur[0]).norm() < ASSERT_TOL
    assert (rdm1_ur[1] - rdm2_ur[1]).norm() < ASSERT_TOL
    # a diagonal rdm is a normalized density matrix of two spins 1/2
    for r in rdm0_dr + rdm0_ur:
        assert r.shape == (2, 2, 2, 2)
        assert abs(r.trace() - 1.0) < 1e-12


def test_symmetrized_simple_update_and_ctmrg_schedules(backend, tmp_path):
    h22, h22b = _hamiltonians()
    tau = 1e-3
    su0 = SimpleUpdate.square_lattice_first_neighbor(h22, backend["D"], tau, degen_ratio=0.9999, rcutoff=1e-12)
    su = SymmetrizedSimpleUpdate.from_infinite_temperature(backend["D"], tau, [h22b])
    assert su.n_bonds == 4
    assert su.beta == 0.0
    assert su.tau == tau

    for k, beta in enumerate(backend["betas"]):
        su.evolve(beta)
        su0.evolve(beta)
        assert su.n_bonds == 4
        assert abs(su.beta - sum(backend["betas"][: k + 1])) < 1e-12
        assert su.tau == tau
        _compare_weights(su0, su)
        assert abs(su.logZ - su0.logZ) < 1e-9
        _compare_schedules(su0, su, backend["chi"], backend["ctm_iters"])

        if k == 0:
            savefile = tmp_path / "save_su_test.npz"
            su.save_to_file(savefile)
            su2 = SymmetrizedSimpleUpdate.load_from_file(savefile)
            assert abs(su2.tau - tau) < 1e-14
            assert abs(su2.beta - su.beta) < 1e-14
            assert su2.D == su.D
            assert su2.Dmax == su.Dmax
            assert abs(su2.degen_ratio - su.degen_ratio) < 1e-14
            assert abs(su2.rcutoff - su.rcutoff) < 1e-14
            assert abs(su2.logZ - su.logZ) < 1e-14
            for t, t2 in zip(su.get_tensors(), su2.get_tensors(), strict=True):
                assert (t - t2).norm() < 1e-14

            savefile0 = tmp_path / "save_su0_test.npz"
            su0.save_to_file(savefile0)
            su3 = SimpleUpdate.load_from_file(savefile0)
            assert abs(su3.logZ - su0.logZ) < 1e-14
            for t, t2 in zip(su0.get_tensors(), su3.get_tensors(), strict=True):
                assert (t - t2).norm() < 1e-14
            with pytest.raises(ValueError, match="does not match class"):
                SimpleUpdate.load_from_file(savefile)
