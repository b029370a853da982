"""Finite-temperature simple update on device SymmetricTensors
(simple_update/simple_update.py:187-698, abstract_simple_update.py:4-187).

Only first-neighbour updates are provided (the square-lattice model used to
seed CTMRG). The update bookkeeping (which bond, which tensors, which leg
permutations before each elementary update, second-order Trotter string with
doubled end gates) follows the reference so that tensors, weights and logZ
match it update by update; every tensor operation runs on the device.
"""

import numpy as np

from ..symmetric_tensor.diagonal_tensor import DiagonalTensor
from ..symmetric_tensor.tools import get_symmetric_tensor_type
from .abstract_simple_update import AbstractSimpleUpdate
from .simple_update_tools import check_hamiltonians, check_tensor_bond_indices


def _leg_swap(init, end):
    return [init.index(v) for v in end]


def update_schedule(tensor_bond_indices, raw_update_data, raw_hamilts, *, second_order=True):
    """Per-update bonds, tensor indices, leg permutations, gate list and the
    initial/final swaps (simple_update.py:11-184, first-neighbour case)."""
    data = np.asarray(raw_update_data, dtype=int)
    if second_order:
        data = np.vstack((data, data[-2:0:-1]))
    bonds = [int(b) for b in data[:, 0]]
    gate_idx = [int(g) for g in data[:, 2]]
    left_idx = [int(i) for i in data[:, 3]]
    right_idx = [int(i) for i in data[:, 4]]
    if np.any(data[:, 0] != data[:, 1]):
        raise NotImplementedError("only first-neighbour updates are implemented")
    n_updates = len(bonds)
    n_tensors = len(tensor_bond_indices)
    hamiltonians = list(raw_hamilts)

    # legs "active" (physical marker -1, updated bond) after a full cycle
    eff = [None] * n_tensors
    for i in range(n_updates):
        eff[left_idx[i]] = [-1, bonds[i]]
        eff[right_idx[i]] = [-1, bonds[i]]
    initial_swap, state = [], []
    for i in range(n_tensors):
        init = [-1, -2, *[int(b) for b in tensor_bond_indices[i]]]
        end = sorted(set(init) - set(eff[i])) + eff[i]
        sw = _leg_swap(init, end)
        initial_swap.append((tuple(sw[:-2]), tuple(sw[-2:])))
        state.append(end)
    lperm, rperm = [], []
    for i in range(n_updates):
        for idx, out in ((left_idx[i], lperm), (right_idx[i], rperm)):
            init = state[idx]
            end = [*sorted(set(init) - {-1, bonds[i]}), -1, bonds[i]]
            sw = _leg_swap(init, end)
            state[idx] = end
            out.append((tuple(sw[:-2]), tuple(sw[-2:])))
    for x in (bonds, gate_idx, left_idx, right_idx, lperm, rperm):
        x.append(x[0])
    if second_order:
        eff[left_idx[0]] = [-1, bonds[0]]
        eff[right_idx[0]] = [-1, bonds[0]]
        hamiltonians.append(2 * raw_hamilts[gate_idx[0]])
        gate_idx[0] = len(raw_hamilts)
        hamiltonians.append(2 * raw_hamilts[gate_idx[n_updates // 2]])
        gate_idx[n_updates // 2] = len(raw_hamilts) + 1
    final_swap = []
    for i in range(n_tensors):
        end = [-1, -2, *[int(b) for b in tensor_bond_indices[i]]]
        init = sorted(set(end) - set(eff[i])) + eff[i]
        sw = _leg_swap(init, end)
        final_swap.append((tuple(sw[:2]), tuple(sw[2:])))
    return bonds, gate_idx, left_idx, right_idx, lperm, rperm, hamiltonians, initial_swap, final_swap


class SimpleUpdate(AbstractSimpleUpdate):
    """Simple update of a thermal ensemble (purification with an ancilla leg)."""

    _classname = "SimpleUpdate"

    def __init__(self, D, beta, tau, rcutoff, degen_ratio, tensor_bond_indices, tensors,
                 raw_update_data, raw_hamilts, weights, logZ, *, verbosity=0):
        self.verbosity = int(verbosity)
        self._tensor_bond_indices = [np.array(t, dtype=int) for t in tensor_bond_indices]
        check_tensor_bond_indices(self._tensor_bond_indices)
        check_hamiltonians(raw_hamilts)
        if len(tensors) != len(self._tensor_bond_indices):
            raise ValueError("Invalid tensor number")
        (
            self._bonds, self._gate_indices, self._left_indices, self._right_indices,
            self._lperm, self._rperm, self._hamiltonians, self._initial_swap, self._final_swap,
        ) = update_schedule(self._tensor_bond_indices, raw_update_data, raw_hamilts)
        self._beta = float(beta)
        self._n_bonds = len(weights)
        self._n_updates = len(self._bonds) - 1
        self._n_tensors = len(tensors)
        self._symmetry = raw_hamilts[0].symmetry()
        self._tensors = list(tensors)
        self._weights = list(weights)
        self.D = int(D)
        self._logZ = float(logZ)
        self.rcutoff = float(rcutoff)
        self.degen_ratio = float(degen_ratio)
        self._raw_update_data = np.asarray(raw_update_data, dtype=int)
        self._raw_hamilts = list(raw_hamilts)
        self.tau = float(tau)

    def __repr__(self):
        return f"SimpleUpdate with {self._symmetry} symmetry and D = {self.D} at beta = {self._beta:.10g}"

    @classmethod
    def square_lattice_first_neighbor(cls, h1, D, tau, *, rcutoff=1e-10, degen_ratio=1.0, verbosity=0):
        """Two-site (AB) square lattice, four bonds (simple_update.py:337-363)."""
        tensor_bond_indices = [[0, 1, 2, 3], [2, 3, 0, 1]]
        update_data = np.array([[b, b, 0, 0, 1, -1] for b in range(4)])
        return cls.from_infinite_temperature(
            D, tau, rcutoff, degen_ratio, tensor_bond_indices, update_data, [h1], verbosity=verbosity
        )

    @classmethod
    def from_infinite_temperature(cls, D, tau, rcutoff, degen_ratio, tensor_bond_indices,
                                  raw_update_data, raw_hamilts, *, verbosity=0):
        """beta = 0 product state: identity between physical and ancilla legs,
        dummy virtual legs (simple_update.py:409-535)."""
        n_tensors = len(tensor_bond_indices)
        n_bonds = max(max(t) for t in tensor_bond_indices) + 1
        ST = type(raw_hamilts[0])
        sched = update_schedule(tensor_bond_indices, raw_update_data, raw_hamilts)
        bonds, gate_idx, left_idx, right_idx = sched[:4]
        hamiltonians = sched[6]
        phys = [ST.singlet() for _ in range(n_tensors)]
        sigs = [[None] * (2 + len(t)) for t in tensor_bond_indices]
        for j, b in enumerate(bonds):
            h = hamiltonians[gate_idx[j]]
            iL, iR = left_idx[j], right_idx[j]
            phys[iL] = h.row_reps[0]
            sigs[iL][0] = ~h.signature[0]
            sigs[iL][1] = h.signature[0]
            sigs[iL][2 + list(tensor_bond_indices[iL]).index(b)] = True
            phys[iR] = h.row_reps[1]
            sigs[iR][0] = ~h.signature[1]
            sigs[iR][1] = h.signature[1]
            sigs[iR][2 + list(tensor_bond_indices[iR]).index(b)] = False
        tensors = []
        for i in range(n_tensors):
            d = ST.representation_dimension(phys[i])
            mat = np.eye(d).reshape((d, d) + (1,) * len(tensor_bond_indices[i]))
            t = ST.from_array(mat, (phys[i], phys[i]), (ST.singlet(),) * len(tensor_bond_indices[i]),
                              signature=sigs[i])
            tensors.append(t)
        weights = [
            DiagonalTensor([np.ones((1,))], ST.singlet(), t.block_irreps, [1], ST.symmetry())
            for _ in range(n_bonds)
        ]
        return cls(D, 0.0, tau, rcutoff, degen_ratio, tensor_bond_indices, tensors, raw_update_data,
                   raw_hamilts, weights, 0.0, verbosity=verbosity)

    @classmethod
    def load_from_file(cls, savefile, *, verbosity=0):
        """Restore from a save_to_file npz (simple_update.py:538-597)."""
        with np.load(savefile) as fin:
            cls._check_savefile(fin, savefile)
            ST = get_symmetric_tensor_type(str(fin["_SimpleUpdate_symmetry"][()]))
            raw_hamilts = [
                ST.load_from_dic(fin, prefix=f"_SimpleUpdate_hamiltonian_{i}")
                for i in range(int(fin["_SimpleUpdate_n_hamiltonians"]))
            ]
            n_tensors = int(fin["_SimpleUpdate_n_tensors"])
            tensors = [ST.load_from_dic(fin, prefix=f"_SimpleUpdate_tensor_{i}") for i in range(n_tensors)]
            tbi = [fin[f"_SimpleUpdate_tensor_bond_indices_{i}"] for i in range(n_tensors)]
            weights = [
                DiagonalTensor.load_from_dic(fin, prefix=f"_SimpleUpdate_weights_{i}")
                for i in range(int(fin["_SimpleUpdate_n_bonds"]))
            ]
            return cls(
                fin["_SimpleUpdate_D"][()], fin["_SimpleUpdate_beta"][()], fin["_SimpleUpdate_tau"][()],
                fin["_SimpleUpdate_rcutoff"][()], fin["_SimpleUpdate_degen_ratio"][()], tbi, tensors,
                fin["_SimpleUpdate_raw_update_data"], raw_hamilts, weights, fin["_SimpleUpdate_logZ"][()],
                verbosity=verbosity,
            )

    def _elementary_update(self, i):
        left = self._tensors[self._left_indices[i]].permute(*self._lperm[i])
        right = self._tensors[self._right_indices[i]].permute(*self._rperm[i])
        b = self._bonds[i]
        left, right, nw, lognf = self.update_first_neighbor(
            left, right, self._weights[b], self._gates[self._gate_indices[i]]
        )
        self._logZ += lognf
        self._weights[b] = nw
        self._tensors[self._left_indices[i]] = left
        self._tensors[self._right_indices[i]] = right

    def update_first_neighbor(self, left, right, weights, gate):
        """Split off the constant parts, apply the gate on theta, truncated SVD back
        to D, normalize the new weights (simple_update.py:633-698)."""
        cstL, effL = left.qr()
        cstR, effR = right.qr()
        effL = effL.permute((0, 1), (2,)) * weights**-1
        effR = effR.permute((2,), (0, 1))
        theta = (effL @ effR).permute((0, 2), (1, 3)) @ gate
        theta = theta.permute((0, 2), (1, 3))
        effL, new_weights, effR = theta.truncated_svd(self.D, rtol=self.rcutoff, degen_ratio=self.degen_ratio)
        new_weights /= new_weights.sum()
        lognf = np.log(theta.norm() / new_weights.norm())
        effL = (effL * new_weights).permute((0,), (1, 2))
        effR = (new_weights * effR).permute((1,), (2, 0))
        return cstL @ effL, cstR @ effR, new_weights, lognf
