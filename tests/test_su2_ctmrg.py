"""SU(2) simple update and CTMRG against their U(1) counterparts (cross-symmetry
invariants of examples/heisenberg_simple_update.py and test_RVB.py of frostspin):
with no binding truncation the SU(2) and U(1) runs are the same physical state,
so weights, logZ, bond representations and energies agree to rounding."""

import numpy as np
import pytest

from frostspin_amd.ctmrg import SequentialCTMRG
from frostspin_amd.ctmrg import models
from frostspin_amd.simple_update import SimpleUpdate

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


def test_su2_simple_update_and_ctmrg_match_u1(backend):
    hS = models.heisenberg_bond("SU2")
    hU = models.heisenberg_bond("U1")
    D, beta = 4, 0.2
    suS = SimpleUpdate.square_lattice_first_neighbor(hS, D, 0.01)
    suU = SimpleUpdate.square_lattice_first_neighbor(hU, D, 0.01)
    suS.evolve(beta)
    suU.evolve(beta)
    # spin 0 + spin 1 <=> U(1) charges {-2, 0, 0, 2} (test_simple_update.py:51-58)
    for rS, rU in zip(suS.get_bond_representations(), suU.get_bond_representations(), strict=True):
        assert rS.tolist() == [[1, 1], [1, 3]]
        assert sorted(rU.tolist()) == [-2, 0, 0, 2]
    for wS, wU in zip(suS.get_weights(), suU.get_weights(), strict=True):
        assert np.abs(wS - wU).max() < 1e-12
    assert abs(suS.logZ - suU.logZ) < 1e-12

    params = {"degen_ratio": 0.99, "cutoff": 1e-10}
    tensors = suS.get_tensors()
    ctmS = SequentialCTMRG.from_elementary_tensors("AB\nBA", tensors, 16, **params)
    ctmU = SequentialCTMRG.from_elementary_tensors("AB\nBA", tensors, 16, **params)
    ctmU.set_symmetry("U1")
    for _ in range(4):
        ctmS.iterate()
        ctmU.iterate()
    assert ctmS.chi_values == ctmU.chi_values
    eS = models.bond_energy(ctmS, hS)
    eU = models.bond_energy(ctmU, hU)
    assert abs(eS - eU) < 1e-12
    assert -0.75 < eS < 0.0


@pytest.mark.gpu
def test_concurrent_site_chains_give_identical_results(gpu):
    """The enlarged corners and renormalizations of the two sites of a move run
    on separate streams; the same iterations with one stream give the same bits
    (energy, every corner and edge tensor)."""
    from frostspin_amd import _ops

    h = models.heisenberg_bond("SU2")
    runs = []
    for enabled in (True, False):
        _ops.CONCURRENCY["enabled"] = enabled
        try:
            _, ctm = models.thermal_ctmrg(h, 6, 1.0, 48)
            for _ in range(5):
                ctm.iterate()
            env = ctm._env
            tensors = [g(x, y) for x, y in ctm._site_coords
                       for g in (env.get_C1, env.get_C2, env.get_C3, env.get_C4,
                                 env.get_T1, env.get_T2, env.get_T3, env.get_T4)]
            runs.append((models.bond_energy(ctm, h), list(ctm.chi_values),
                         [_ops.download(t._data) for t in tensors]))
        finally:
            _ops.CONCURRENCY["enabled"] = False
    (e1, chi1, d1), (e2, chi2, d2) = runs
    assert chi1 == chi2 and e1 == e2
    for a, b in zip(d1, d2, strict=True):
        assert np.array_equal(a, b)
