"""Spin-1/2 bond Hamiltonians as symmetric tensors and the thermal-state setup
used to seed CTMRG (simple update from beta = 0), following
examples/heisenberg_simple_update.py:22-111 of frostspin.
"""

import numpy as np

from ..simple_update.simple_update import SimpleUpdate
from ..symmetric_tensor.su2_symmetric_tensor import SU2SymmetricTensor
from ..symmetric_tensor.u1_symmetric_tensor import U1SymmetricTensor
from .sequential_ctmrg import SequentialCTMRG

SPIN_HALF_U1 = np.array([1, -1], dtype=np.int8)  # 2 S_z of |up>, |down>
SPIN_HALF_SU2 = np.array([[1], [2]])  # one spin-1/2 multiplet

# S.S on two spins 1/2 (examples/heisenberg_simple_update.py:25-32)
SDS_22 = np.array(
    [
        [0.25, 0.0, 0.0, 0.0],
        [0.0, -0.25, 0.5, 0.0],
        [0.0, 0.5, -0.25, 0.0],
        [0.0, 0.0, 0.0, 0.25],
    ]
).reshape(2, 2, 2, 2)


def heisenberg_bond(symmetry="SU2"):
    """h = S.S as a (2, 2 | 2, 2) SU(2) or U(1) tensor."""
    if symmetry == "SU2":
        reps = (SPIN_HALF_SU2, SPIN_HALF_SU2)
        return SU2SymmetricTensor.from_array(SDS_22, reps, reps)
    if symmetry == "U1":
        return xxz_bond(1.0, 1.0)
    raise ValueError(f"unsupported symmetry {symmetry}")


def xxz_bond(jz=1.0, jperp=1.0):
    """h = Jperp (Sx Sx + Sy Sy) + Jz Sz Sz on two spins 1/2, as a (2, 2 | 2, 2)
    U(1) tensor (Heisenberg for Jz = Jperp = 1, test_simple_update.py:16-26)."""
    h = np.array(
        [
            [jz / 4, 0.0, 0.0, 0.0],
            [0.0, -jz / 4, jperp / 2, 0.0],
            [0.0, jperp / 2, -jz / 4, 0.0],
            [0.0, 0.0, 0.0, jz / 4],
        ]
    ).reshape(2, 2, 2, 2)
    reps = (SPIN_HALF_U1, SPIN_HALF_U1)
    return U1SymmetricTensor.from_array(h, reps, reps)


def thermal_tensors(h, D, beta, *, tau=0.01, rcutoff=1e-10, degen_ratio=0.99):
    """Simple update from infinite temperature to beta; returns the two site
    tensors of the AB/BA cell with sqrt(weights) absorbed."""
    su = SimpleUpdate.square_lattice_first_neighbor(h, D, tau, rcutoff=rcutoff, degen_ratio=degen_ratio)
    su.evolve(beta)
    return su, su.get_tensors()


def thermal_ctmrg(h, D, beta, chi, *, tau=0.01, cutoff=1e-10, degen_ratio=0.99):
    """SequentialCTMRG on the AB/BA cell seeded by simple update at (D, beta)."""
    su, tensors = thermal_tensors(h, D, beta, tau=tau)
    ctm = SequentialCTMRG.from_elementary_tensors(
        "AB\nBA", tensors, chi, cutoff=cutoff, degen_ratio=degen_ratio
    )
    return su, ctm


def bond_energy(ctm, h):
    """Average <h> over the horizontal and vertical bonds of the cell, from the
    1x2 and 2x1 reduced density matrices (examples/heisenberg_simple_update.py)."""
    r12, r21 = ctm.compute_rdm_1st_neighbor_cell()
    hd = h.toarray().reshape(4, 4)
    e = [float(np.trace(r.toarray().reshape(4, 4) @ hd)) for r in r12 + r21]
    return float(np.mean(e))
