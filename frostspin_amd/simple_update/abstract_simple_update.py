"""Shared simple-update machinery (simple_update/abstract_simple_update.py:5-187).

Properties, gate construction, the second-order Trotter string of evolve()
and file I/O in the reference's npz key layout, so save files interchange.
Subclasses supply the update schedule (_elementary_update, _initial_swap,
_final_swap, _n_updates) and load_from_file.
"""

import numpy as np


class AbstractSimpleUpdate:
    _is_second_order = True

    @property
    def Dmax(self):
        return max(max(t.shape[2:]) for t in self._tensors)

    @property
    def beta(self):
        return self._beta

    @property
    def n_bonds(self):
        return self._n_bonds

    @property
    def n_tensors(self):
        return self._n_tensors

    @property
    def logZ(self):
        return self._logZ / self._n_tensors

    @property
    def tau(self):
        return self._tau

    @tau.setter
    def tau(self, tau):
        if self.verbosity > 0:
            print(f"set tau to {tau}")
        self._tau = tau
        self._gates = tuple((-tau * h).expm() for h in self._hamiltonians)
        # 2nd order Trotter-Suzuki and rho quadratic in psi (abstract_simple_update.py:29-36)
        self._dbeta = 2 * (1 + self._is_second_order) * tau

    def __str__(self):
        return (
            f"{self!r}\nn_tensors = {self._n_tensors}, n_bonds = {self._n_bonds}"
            f"\nDmax = {self.Dmax}, tau = {self._tau}, rcutoff = {self.rcutoff}, "
            f"degen_ratio = {self.degen_ratio}"
        )

    def symmetry(self):
        return self._symmetry

    def get_bond_representation(self, i):
        for j in range(self._n_tensors):
            ind = (np.asarray(self._tensor_bond_indices[j]) == i).nonzero()[0]
            if ind.size:
                return self._tensors[j].col_reps[ind[0]]
        raise ValueError(f"Unknown bond {i}")

    def get_bond_representations(self):
        return [self.get_bond_representation(i) for i in range(self._n_bonds)]

    def get_weights(self, *, sort=True):
        return [w.toarray(sort=sort) for w in self._weights]

    def bond_entanglement_entropy(self):
        return np.array([-w @ np.log(w) for w in self.get_weights()])

    def get_tensors(self):
        """Tensors with sqrt(weights) absorbed on every virtual leg
        (abstract_simple_update.py:73-92)."""
        sqw = [w**-0.5 for w in self._weights]
        out = []
        for i, t0 in enumerate(self._tensors):
            rswap = (0, 1, *range(3, t0.ndim))
            t = t0
            for leg in self._tensor_bond_indices[i]:
                t = t.permute(rswap, (2,)) * sqw[leg]
            out.append(t.permute((0, 1), tuple(range(2, t0.ndim))))
        return out

    def evolve(self, beta_evolve):
        """Imaginary-time evolution by beta_evolve with the 2nd-order Trotter string
        (abstract_simple_update.py:94-147): the first and last half-step updates
        are run once with tau, the ones in between merge into 2 tau updates."""
        if self.verbosity > 0:
            print(f"Evolve in imaginary time for beta from {self._beta:.6g} to {self._beta + beta_evolve:.6g}...")
        if beta_evolve < -1e-12:
            raise ValueError("Cannot evolve for negative imaginary time")
        if beta_evolve < 0.9 * self._dbeta:
            return
        niter = round(beta_evolve / self._dbeta)
        for i in range(self._n_tensors):
            self._tensors[i] = self._tensors[i].permute(*self._initial_swap[i])
        self._elementary_update(self._n_updates)
        for j in range(1, self._n_updates):
            self._elementary_update(j)
        for _ in range(niter - 1):
            for j in range(self._n_updates):
                self._elementary_update(j)
        if self._is_second_order:
            self._elementary_update(self._n_updates)
        for i in range(self._n_tensors):
            self._tensors[i] = self._tensors[i].permute(*self._final_swap[i])
        self._beta += niter * self._dbeta

    def save_to_file(self, savefile, **additional_data):
        """npz with the reference's keys (abstract_simple_update.py:149-187)."""
        data = {
            "_SimpleUpdate_symmetry": self._symmetry,
            "_SimpleUpdate_classname": self._classname,
            "_SimpleUpdate_D": self.D,
            "_SimpleUpdate_beta": self._beta,
            "_SimpleUpdate_tau": self._tau,
            "_SimpleUpdate_logZ": self._logZ,
            "_SimpleUpdate_is_second_order": self._is_second_order,
            "_SimpleUpdate_rcutoff": self.rcutoff,
            "_SimpleUpdate_degen_ratio": self.degen_ratio,
            "_SimpleUpdate_raw_update_data": self._raw_update_data,
            "_SimpleUpdate_n_bonds": self._n_bonds,
            "_SimpleUpdate_n_hamiltonians": len(self._raw_hamilts),
            "_SimpleUpdate_n_tensors": self._n_tensors,
        }
        for i, h in enumerate(self._raw_hamilts):
            data |= h.get_data_dic(prefix=f"_SimpleUpdate_hamiltonian_{i}")
        for i, tbi in enumerate(self._tensor_bond_indices):
            data |= self._tensors[i].get_data_dic(prefix=f"_SimpleUpdate_tensor_{i}")
            data[f"_SimpleUpdate_tensor_bond_indices_{i}"] = np.asarray(tbi)
        for i, w in enumerate(self._weights):
            data |= w.get_data_dic(prefix=f"_SimpleUpdate_weights_{i}")
        np.savez_compressed(savefile, **data, **additional_data)
        if self.verbosity > 0:
            print("Simple update saved in file", savefile)

    @classmethod
    def _check_savefile(cls, fin, savefile):
        if fin["_SimpleUpdate_classname"] != cls._classname:
            raise ValueError(f"Savefile '{savefile}' does not match class '{cls._classname}'")
