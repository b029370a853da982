"""Argument validation of the simple-update classes
(simple_update/simple_update_tools.py:4-42); same rules, same ValueErrors."""

import numpy as np


def check_tensor_bond_indices(tensor_bond_indices):
    """Bonds are labelled 0 .. n_bonds-1, each appears on exactly two tensors
    and at most once per tensor."""
    n_bonds = max(max(tbi) for tbi in tensor_bond_indices) + 1
    seen = np.zeros((n_bonds,), dtype=int)
    for i, tbi in enumerate(tensor_bond_indices):
        legs = [int(v) for v in tbi]
        for leg in legs:
            if not 0 <= leg < n_bonds:
                raise ValueError("Bond indices must be 0 to n_bonds - 1")
        dup = {v for v in legs if legs.count(v) > 1}
        if dup:
            raise ValueError(f"Leg {min(dup)} apperars twice in tensor {i}")
        for leg in legs:
            seen[leg] += 1
    bad = (seen != 2).nonzero()[0]
    if bad.size:
        raise ValueError(f"Virtual bond {bad[0]} appears {seen[bad[0]]} times")


def check_hamiltonians(hamiltonians):
    """Two-site operators of one tensor type: rank 4 split 2|2, equal row and
    column representations, opposite signatures on each site."""
    ST = type(hamiltonians[0])
    for i, h in enumerate(hamiltonians):
        if type(h) is not ST:
            raise ValueError(f"Invalid type for Hamiltonian {i}")
        if h.ndim != 4 or h.n_row_reps != 2:
            raise ValueError(f"Hamiltonian {i} has invalid shape")
        for a in range(2):
            r, c = np.asarray(h.row_reps[a]), np.asarray(h.col_reps[a])
            if r.shape != c.shape or (r != c).any():
                raise ValueError(f"Hamiltonian {i} has invalid representations")
            if not h.signature[a] ^ h.signature[a + 2]:
                raise ValueError(f"Hamiltonian {i} has invalid signature")
