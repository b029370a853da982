"""Simple update of an AB/BA square-lattice state whose B tensor is the
rotated conjugate of A (simple_update/symmetrized_simple_update.py:9-271).

Only A is stored. Every bond update contracts A with its own conjugate:
theta = effL effL^dagger, the gate acts on both physical legs, and the
truncated eigen-decomposition of the (real symmetric) theta gives the new
weights and the new A half at once. The bond / gate schedule and the leg
permutations of each elementary update follow the reference, so tensors,
weights and logZ match it update by update.
"""

import numpy as np

from ..symmetric_tensor.diagonal_tensor import DiagonalTensor
from ..symmetric_tensor.tools import get_symmetric_tensor_type
from .abstract_simple_update import AbstractSimpleUpdate
from .simple_update_tools import check_hamiltonians


class SymmetrizedSimpleUpdate(AbstractSimpleUpdate):
    # leg layout of A outside evolve: (phys, anc, bond0, bond1, bond2, bond3)
    _initial_swap = [((1, 2, 4, 5), (3, 0))]
    _final_swap = [((5, 0), (4, 1, 2, 3))]
    # per elementary update: permute A to (constant legs..., updated bond, phys)
    _lperm = [
        ((0, 4, 2, 3), (1, 5)),
        ((0, 4, 2, 3), (1, 5)),
        ((0, 1, 4, 3), (2, 5)),
        ((0, 1, 2, 4), (3, 5)),
        ((0, 1, 2, 4), (3, 5)),
        ((0, 1, 4, 3), (2, 5)),
        ((0, 4, 2, 3), (1, 5)),
    ]
    _tensor_bond_indices = [np.array([0, 1, 2, 3])]
    _1st_updated_bond = [0, 1, 2, 3, 2, 1, 0]
    _gate_indices = [1, 0, 0, 2, 0, 0, 0]
    _raw_update_data = np.zeros((0, 0), dtype=int)
    _n_bonds = 4
    _n_updates = 6
    _n_tensors = 1
    _classname = "SymmetrizedSimpleUpdate"

    def __init__(self, D, beta, tau, rcutoff, degen_ratio, A, raw_hamilts, weights, logZ, *, verbosity=0):
        self.verbosity = int(verbosity)
        if len(weights) != 4:
            raise ValueError("Invalid weight length")
        check_hamiltonians(raw_hamilts)
        self._beta = float(beta)
        self._symmetry = raw_hamilts[0].symmetry()
        self._tensors = [A]
        self._weights = list(weights)
        self.D = int(D)
        self._logZ = float(logZ)
        self.rcutoff = float(rcutoff)
        self.degen_ratio = float(degen_ratio)
        self._raw_hamilts = list(raw_hamilts)
        h = self._raw_hamilts
        if len(h) == 1:
            self._hamiltonians = [h[0], 2 * h[0], 2 * h[0]]
        elif len(h) == 4:
            self._hamiltonians = [h[0], h[1], h[2], 2 * h[3], 2 * h[0]]
        else:
            raise ValueError("Invalid raw_hamiltonians")
        self.tau = float(tau)

    def __repr__(self):
        return f"SymmetrizedSimpleUpdate with {self._symmetry} symmetry and D = {self.D} at beta = {self._beta:.10g}"

    @property
    def logZ(self):
        # two sites per cell although one tensor is stored
        return self._logZ / 2

    @classmethod
    def from_infinite_temperature(cls, D, tau, raw_hamilts, *, rcutoff=1e-14, degen_ratio=0.99999, verbosity=0):
        """beta = 0 product state (symmetrized_simple_update.py:124-183)."""
        ST = type(raw_hamilts[0])
        phys_rep = raw_hamilts[0].row_reps[0]
        d = ST.representation_dimension(phys_rep)
        if not raw_hamilts[0].signature[0] ^ raw_hamilts[0].signature[1]:
            raise ValueError("Hamiltonian sites must carry opposite signatures")
        s0 = bool(raw_hamilts[0].signature[0])
        a0 = ST.from_array(
            np.eye(d).reshape((d, d, 1, 1, 1, 1)), (phys_rep, phys_rep), (ST.singlet(),) * 4,
            signature=[not s0, s0, True, True, True, True],
        )
        weights = [
            DiagonalTensor([np.ones((1,))], ST.singlet(), a0.block_irreps, [1], ST.symmetry()) for _ in range(4)
        ]
        return cls(D, 0.0, tau, rcutoff, degen_ratio, a0, raw_hamilts, weights, 0.0, verbosity=verbosity)

    def get_tensors(self):
        """A with sqrt(weights) on its virtual legs, and B = rotated A^dagger."""
        t = self._tensors[0]
        for i in range(4):
            t = t.permute((0, 1, 3, 4, 5), (2,)) * self._weights[i] ** -0.5
        t = t.permute((0, 1), (2, 3, 4, 5))
        return [t, t.dagger().permute((4, 5), (2, 3, 0, 1))]

    def update_first_neighbor(self, left, weights, gate):
        """(symmetrized_simple_update.py:207-229)"""
        cstL, effL = left.qr()
        effL = effL.permute((0, 2), (1,)) * weights ** (-1 / 2)
        theta = (effL @ effL.dagger()).permute((0, 2), (1, 3)) @ gate
        theta = theta.permute((0, 2), (1, 3))
        new_weights, new_effL = theta.truncated_eigh(self.D, rtol=self.rcutoff, degen_ratio=self.degen_ratio)
        new_weights /= new_weights.sum()
        lognf = np.log(theta.norm() / new_weights.norm())
        new_effL = (new_effL * new_weights).permute((0,), (2, 1))
        return cstL @ new_effL, new_weights, lognf

    def _elementary_update(self, i):
        left = self._tensors[0].permute(*self._lperm[i])
        b = self._1st_updated_bond[i]
        new_left, nw, lognf = self.update_first_neighbor(left, self._weights[b], self._gates[self._gate_indices[i]])
        self._logZ += lognf
        self._weights[b] = nw
        self._tensors[0] = new_left

    @classmethod
    def load_from_file(cls, savefile, *, verbosity=0):
        """(symmetrized_simple_update.py:231-271)"""
        with np.load(savefile) as fin:
            cls._check_savefile(fin, savefile)
            ST = get_symmetric_tensor_type(str(fin["_SimpleUpdate_symmetry"][()]))
            raw_hamilts = [
                ST.load_from_dic(fin, prefix=f"_SimpleUpdate_hamiltonian_{i}")
                for i in range(int(fin["_SimpleUpdate_n_hamiltonians"]))
            ]
            A = ST.load_from_dic(fin, prefix="_SimpleUpdate_tensor_0")
            weights = [DiagonalTensor.load_from_dic(fin, prefix=f"_SimpleUpdate_weights_{i}") for i in range(4)]
            return cls(
                fin["_SimpleUpdate_D"][()], fin["_SimpleUpdate_beta"][()], fin["_SimpleUpdate_tau"][()],
                fin["_SimpleUpdate_rcutoff"][()], fin["_SimpleUpdate_degen_ratio"][()], A, raw_hamilts,
                weights, fin["_SimpleUpdate_logZ"][()], verbosity=verbosity,
            )
