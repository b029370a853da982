"""State, Trotter schedule and checkpoint of a simple update
(interface of simple_update/abstract_simple_update.py:5-187).

A simple update holds n_tensors site tensors and n_bonds diagonal weights. One
call of evolve(b) runs round(b / dbeta) Trotter steps, dbeta = 2 (1 + second
order) tau (the density matrix is quadratic in the tensors). In the
second-order string the first gate of a cycle is applied with tau on both ends
of the whole evolution and merged into 2 tau gates in between, so the update
indices run

    [n_updates (tau), 1 .. n_updates - 1, (0 .. n_updates - 1) x (steps - 1),
     n_updates (tau)]

and the tensors are permuted into the leg order the cycle expects before the
string and back after it. That order and the npz keys are those of the
reference, so tensors, weights and checkpoints interchange. Subclasses supply
_elementary_update, _initial_swap, _final_swap, _n_updates, _classname and
load_from_file.
"""

import numpy as np

# checkpoint keys: npz key suffix -> attribute holding the value
_SAVED_SCALARS = (
    ("symmetry", "_symmetry"),
    ("classname", "_classname"),
    ("D", "D"),
    ("beta", "_beta"),
    ("tau", "_tau"),
    ("logZ", "_logZ"),
    ("is_second_order", "_is_second_order"),
    ("rcutoff", "rcutoff"),
    ("degen_ratio", "degen_ratio"),
    ("raw_update_data", "_raw_update_data"),
    ("n_bonds", "_n_bonds"),
    ("n_tensors", "_n_tensors"),
)
_KEY = "_SimpleUpdate_"


def _read_only(attr, doc):
    return property(lambda self: getattr(self, attr), doc=doc)


class AbstractSimpleUpdate:
    _is_second_order = True

    beta = _read_only("_beta", "Inverse temperature reached so far.")
    n_bonds = _read_only("_n_bonds", "Number of distinct bonds (weights).")
    n_tensors = _read_only("_n_tensors", "Number of stored site tensors.")

    @property
    def Dmax(self):
        """Largest virtual leg dimension over the site tensors."""
        return max(d for t in self._tensors for d in t.shape[2:])

    @property
    def logZ(self):
        """Accumulated log of the normalization factors, per site tensor."""
        return self._logZ / self._n_tensors

    @property
    def tau(self):
        return self._tau

    @tau.setter
    def tau(self, tau):
        """New time step: gates exp(-tau h) rebuilt on the device."""
        if self.verbosity > 0:
            print(f"simple update: time step tau = {tau}")
        self._tau = tau
        self._gates = tuple((-tau * h).expm() for h in self._hamiltonians)
        self._dbeta = 2 * (1 + self._is_second_order) * tau

    def __str__(self):
        return (
            f"{self!r}: {self._n_tensors} tensors, {self._n_bonds} bonds, Dmax {self.Dmax}, "
            f"tau {self._tau}, rcutoff {self.rcutoff}, degen_ratio {self.degen_ratio}"
        )

    def symmetry(self):
        return self._symmetry

    # bonds ----------------------------------------------------------------------
    def _bond_owner(self, bond):
        """(tensor index, virtual leg index) of the first tensor carrying bond."""
        for i, legs in enumerate(self._tensor_bond_indices):
            hits = np.flatnonzero(np.asarray(legs) == bond)
            if hits.size:
                return i, int(hits[0])
        raise ValueError(f"bond {bond} is not attached to any tensor")

    def get_bond_representation(self, i):
        t, leg = self._bond_owner(i)
        return self._tensors[t].col_reps[leg]

    def get_bond_representations(self):
        return [self.get_bond_representation(b) for b in range(self._n_bonds)]

    def get_weights(self, *, sort=True):
        return [w.toarray(sort=sort) for w in self._weights]

    def bond_entanglement_entropy(self):
        return np.array([-(w @ np.log(w)) for w in self.get_weights()])

    def get_tensors(self):
        """Site tensors with sqrt(weight) on every virtual leg
        (abstract_simple_update.py:73-92). Between updates both ends of a bond
        carry its full weight w, so each leg, brought to the single column in
        turn, is scaled by w^-1/2."""
        inv_sqrt = [w**-0.5 for w in self._weights]
        out = []
        for t, legs in zip(self._tensors, self._tensor_bond_indices, strict=True):
            rest = (0, 1, *range(3, t.ndim))  # rotates the virtual legs by one
            for bond in legs:
                t = t.permute(rest, (2,)) * inv_sqrt[bond]
            out.append(t.permute((0, 1), tuple(range(2, t.ndim))))
        return out

    # imaginary-time evolution ---------------------------------------------------------
    def _trotter_string(self, steps):
        """Update indices of `steps` Trotter steps (module docstring)."""
        n = self._n_updates
        yield n
        yield from range(1, n)
        for _ in range(steps - 1):
            yield from range(n)
        if self._is_second_order:
            yield n

    def evolve(self, beta_evolve):
        """Imaginary-time evolution by beta_evolve (abstract_simple_update.py:94-147)."""
        if beta_evolve < -1e-12:
            raise ValueError("imaginary time cannot decrease")
        if beta_evolve < 0.9 * self._dbeta:  # less than one step (float rounding allowed)
            return
        steps = round(beta_evolve / self._dbeta)
        if self.verbosity > 0:
            print(f"simple update: beta {self._beta:.6g} -> {self._beta + steps * self._dbeta:.6g}")
        self._tensors = [t.permute(*sw) for t, sw in zip(self._tensors, self._initial_swap, strict=True)]
        for j in self._trotter_string(steps):
            self._elementary_update(j)
        self._tensors = [t.permute(*sw) for t, sw in zip(self._tensors, self._final_swap, strict=True)]
        self._beta += steps * self._dbeta

    # checkpoint -------------------------------------------------------------------
    def save_to_file(self, savefile, **additional_data):
        """npz checkpoint in the reference's key layout
        (abstract_simple_update.py:149-187)."""
        data = {_KEY + key: getattr(self, attr) for key, attr in _SAVED_SCALARS}
        data[_KEY + "n_hamiltonians"] = len(self._raw_hamilts)
        for i, h in enumerate(self._raw_hamilts):
            data |= h.get_data_dic(prefix=f"{_KEY}hamiltonian_{i}")
        for i, (t, legs) in enumerate(zip(self._tensors, self._tensor_bond_indices, strict=True)):
            data |= t.get_data_dic(prefix=f"{_KEY}tensor_{i}")
            data[f"{_KEY}tensor_bond_indices_{i}"] = np.asarray(legs)
        for i, w in enumerate(self._weights):
            data |= w.get_data_dic(prefix=f"{_KEY}weights_{i}")
        np.savez_compressed(savefile, **data, **additional_data)
        if self.verbosity > 0:
            print("simple update written to", savefile)

    @classmethod
    def _check_savefile(cls, fin, savefile):
        if fin[_KEY + "classname"] != cls._classname:
            raise ValueError(f"{savefile} holds a {fin[_KEY + 'classname']}: does not match class {cls._classname}")
