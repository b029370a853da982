"""CTMRG environment container (ctmrg/ctm_environment.py:6-460).

Holds, for every inequivalent site of the unit cell, the site tensor A, the four
corners C1..C4, the four edges T1..T4, cached enlarged corners and the
projectors / renormalized tensors of the move in progress. All tensors stay
device-resident between iterations. Coordinates are periodic: (x, y) maps to
the site index indices[x mod Lx, y mod Ly].
"""

import numpy as np

from ..symmetric_tensor.tools import get_symmetric_tensor_type
from .ctm_contract import forget_site_permutes


def initialize_dummy_env(A):
    """Corners = 1x1 identities with trivial representation; edges = identities
    between bra and ket layers with dummy legs (ctm_environment.py:6-34)."""
    ST = type(A)
    one = (ST.singlet(),)
    C1 = ST.from_array(np.eye(1), one, one)
    C2 = ST.from_array(np.eye(1), one, one)
    C3 = ST.from_array(np.eye(1), one, one, signature=[True, False])
    C4 = ST.from_array(np.eye(1), one, one)

    def edge(leg, layout):
        # layout: where the identity sits among the four legs
        d = A.shape[leg]
        rep = A.col_reps[leg - 2]
        s = A.signature[leg]
        if layout == "T1":
            return ST.from_array(np.eye(d)[None, :, :, None], one, (rep, rep, ST.singlet()),
                                 signature=np.array([False, s, ~s, True]))
        if layout == "T2":
            return ST.from_array(np.eye(d)[None, None, :, :], one, (ST.singlet(), rep, rep),
                                 signature=np.array([True, False, s, ~s]))
        if layout == "T3":
            return ST.from_array(np.eye(d)[:, :, None, None], (rep, rep, ST.singlet()), one,
                                 signature=np.array([s, ~s, True, False]))
        return ST.from_array(np.eye(d)[None, :, :, None], one, (rep, rep, ST.singlet()),
                             signature=np.array([False, s, ~s, True]))

    T1 = edge(4, "T1")
    T2 = edge(5, "T2")
    T3 = edge(2, "T3")
    T4 = edge(3, "T4")
    return C1, T1, C2, T2, C3, T3, C4, T4


def find_unique_coords(cell):
    """Site letters in order of first appearance; (x, y) <-> site index maps
    (ctm_environment.py:37-59)."""
    letters = list(dict.fromkeys(cell.flat))
    n_sites = len(letters)
    site_coords = np.empty((n_sites, 2), dtype=np.int8)
    indicesT = np.empty(cell.shape, dtype=int)
    for i, c in enumerate(letters):
        mask = cell == c
        rows, cols = mask.nonzero()
        site_coords[i] = cols[0], rows[0]
        indicesT[mask] = i
    return site_coords, indicesT.T.copy()


def parse_tiling(tiling):
    txt = [" ".join(w) for w in tiling.strip().splitlines()]
    return np.atleast_2d(np.genfromtxt(txt, dtype="U1"))


# tensors kept per inequivalent site, in npz order (ctm_environment.py:239-281)
_SITE_TENSORS = ("A", "C1", "C2", "C3", "C4", "T1", "T2", "T3", "T4")
_EXPECTED_NDIM = {"A": 6, "C1": 2, "C2": 2, "C3": 2, "C4": 2, "T1": 4, "T2": 4, "T3": 4, "T4": 4}
# one move renormalizes a corner, the edge and the next corner on one side
_MOVE_TARGETS = {
    "up": ("C1", "T1", "C2"),
    "right": ("C2", "T2", "C3"),
    "down": ("C3", "T3", "C4"),
    "left": ("C4", "T4", "C1"),
}
_ENLARGED = ("ul", "ur", "dl", "dr")


class CTMEnvironment:
    """Device-resident CTMRG environment of a periodic unit cell (interface and
    leg conventions of ctmrg/ctm_environment.py:62-460).

    Storage: one list per tensor kind (site tensor A, corners C1..C4, edges
    T1..T4) indexed by inequivalent site, plus the per-move scratch (projectors
    of each direction, renormalized corner/edge/corner) and the cached enlarged
    corners. (x, y) is periodic and maps to a site through the cell."""

    def __init__(self, cell, As, C1s, C2s, C3s, C4s, T1s, T2s, T3s, T4s):
        self._cell = cell
        self._site_coords, self._indices = find_unique_coords(cell)
        self._n_sites = self._site_coords.shape[0]
        self._Ly, self._Lx = cell.shape
        self._site = {"A": tuple(As)}
        for name, ts in zip(_SITE_TENSORS[1:], (C1s, C2s, C3s, C4s, T1s, T2s, T3s, T4s), strict=True):
            self._site[name] = list(ts)
        self._symmetry = type(As[0]).symmetry()
        self.check_consistency()
        self.reset_constructed_corners()
        self._clear_move()

    # cell geometry ----------------------------------------------------------------
    Lx = property(lambda self: self._Lx)
    Ly = property(lambda self: self._Ly)
    cell = property(lambda self: self._cell)
    n_sites = property(lambda self: self._n_sites)
    site_coords = property(lambda self: self._site_coords)
    elementary_tensors = property(lambda self: self._site["A"])

    def _unique_index(self, x, y):
        return self._indices[x % self._Lx, y % self._Ly]

    def get_tensor_type(self, x, y):
        return self._cell[x % self._Lx, y % self._Ly]

    def symmetry(self):
        return self._symmetry

    @property
    def Dmax(self):
        return max(d for A in self._site["A"] for d in A.shape[2:])

    @property
    def Dmin(self):
        return min(d for A in self._site["A"] for d in A.shape[2:])

    @property
    def chi_values(self):
        """Distinct corner dimensions over the cell."""
        return sorted({d for C in self._site["C1"] + self._site["C3"] for d in C.shape})

    def get_corner_representations(self):
        """Distinct corner leg representations, by dimension then content."""
        reps = {}
        for C in self._site["C1"] + self._site["C3"]:
            for r in (C.row_reps[0], C.col_reps[0]):
                reps.setdefault(tuple(np.asarray(r).ravel()), r)
        ST = type(self._site["A"][0])
        return sorted(reps.values(), key=lambda r: (ST.representation_dimension(r), tuple(np.asarray(r).ravel())))

    def check_consistency(self):
        """Every kind has one tensor of the right type and rank per site, and
        neighbouring site tensors share bond representations with opposite
        signatures."""
        ST = type(self._site["A"][0])
        for name in _SITE_TENSORS:
            ts = self._site[name]
            if len(ts) != self._n_sites:
                raise ValueError(f"{name}: {len(ts)} tensors for {self._n_sites} sites")
            for t in ts:
                if type(t) is not ST:
                    raise ValueError(f"{name}: {type(t).__name__} among {ST.__name__} tensors")
                if t.ndim != _EXPECTED_NDIM[name]:
                    raise ValueError(f"{name}: rank {t.ndim} instead of {_EXPECTED_NDIM[name]}")
        # (neighbour offset, leg of A, leg of the neighbour)
        for x, y in self._site_coords:
            A = self.get_A(x, y)
            for (dx, dy), leg, nleg in (((0, -1), 2, 4), ((1, 0), 3, 5)):
                B = self.get_A(x + dx, y + dy)
                where = f"sites {(x, y)} and {(x + dx, y + dy)}"
                if A.signature[leg] == B.signature[nleg]:
                    raise ValueError(f"signature mismatch between {where}")
                if not np.array_equal(np.asarray(A.col_reps[leg - 2]), np.asarray(B.col_reps[nleg - 2])):
                    raise ValueError(f"representation mismatch between {where}")

    # construction, I/O -------------------------------------------------------------
    @classmethod
    def from_elementary_tensors(cls, tiling, tensors):
        cell = parse_tiling(tiling)
        if len(tensors) != len(set(cell.flat)):
            raise ValueError("the cell and the number of tensors do not agree")
        per_site = [initialize_dummy_env(A) for A in tensors]  # C1, T1, C2, T2, C3, T3, C4, T4
        C1s, T1s, C2s, T2s, C3s, T3s, C4s, T4s = (list(k) for k in zip(*per_site, strict=True))
        return cls(cell, tensors, C1s, C2s, C3s, C4s, T1s, T2s, T3s, T4s)

    @classmethod
    def load_from_file(cls, savefile):
        """Environment from an npz in the reference's schema (ctm_environment.py:239-281)."""
        with np.load(savefile) as data:
            cell = data["_CTM_cell"]
            ST = get_symmetric_tensor_type(str(data["_CTM_symmetry"][()]))
            n = len(set(cell.flat))
            lists = [[ST.load_from_dic(data, prefix=f"_CTM_{k}_{i}") for i in range(n)] for k in _SITE_TENSORS]
        return cls(cell, *lists)

    def get_data_dic(self):
        data = {"_CTM_cell": self._cell, "_CTM_symmetry": self._symmetry}
        for k in _SITE_TENSORS:
            for i, t in enumerate(self._site[k]):
                data |= t.get_data_dic(prefix=f"_CTM_{k}_{i}")
        return data

    # changing tensors -----------------------------------------------------------------
    def set_tensors(self, tensors):
        if len(tensors) != self._n_sites:
            raise ValueError("the cell and the number of tensors do not agree")
        for new, old in zip(tensors, self._site["A"], strict=True):
            if not new.match_representations(old):
                raise ValueError("a new site tensor does not carry the old representations")
        self._site["A"] = tuple(tensors)
        forget_site_permutes()
        self.reset_constructed_corners()

    def set_symmetry(self, symmetry):
        if symmetry == self._symmetry:
            return
        self.reset_constructed_corners()
        self._symmetry = symmetry
        self._site = {k: (tuple if k == "A" else list)(t.cast(symmetry) for t in ts)
                      for k, ts in self._site.items()}

    # per-move scratch: projectors, renormalized tensors, enlarged corners ------------
    def _clear_move(self):
        n = self._n_sites
        self._projectors = {d: ([None] * n, [None] * n) for d in _MOVE_TARGETS}
        self._renormalized = [[None] * n for _ in range(3)]

    def reset_constructed_corners(self):
        self._enlarged = {c: [None] * self._n_sites for c in _ENLARGED}

    def store_renormalized_tensors(self, x, y, nCX, nT, nCY):
        j = self._unique_index(x, y)
        for lst, t in zip(self._renormalized, (nCX, nT, nCY), strict=True):
            lst[j] = t

    def _commit_move(self, direction):
        for name, lst in zip(_MOVE_TARGETS[direction], self._renormalized, strict=True):
            self._site[name] = lst
        self._clear_move()


def _site_getter(name):
    def get(self, x, y):
        return self._site[name][self._unique_index(x, y)]

    get.__name__ = get.__qualname__ = f"get_{name}"
    return get


def _projector_accessors(direction):
    def store(self, x, y, P, Pt):
        j = self._unique_index(x, y)
        self._projectors[direction][0][j] = P
        self._projectors[direction][1][j] = Pt

    def get_P(self, x, y):
        return self._projectors[direction][0][self._unique_index(x, y)]

    def get_Pt(self, x, y):
        return self._projectors[direction][1][self._unique_index(x, y)]

    def commit(self):
        self._commit_move(direction)

    return {
        f"store_{direction}_projectors": store,
        f"get_{direction}_P": get_P,
        f"get_{direction}_Pt": get_Pt,
        f"set_renormalized_tensors_{direction}": commit,
    }


def _enlarged_accessors(corner):
    def get(self, x, y):
        return self._enlarged[corner][self._unique_index(x, y)]

    def put(self, x, y, value):
        self._enlarged[corner][self._unique_index(x, y)] = value

    return {f"get_corner_{corner}": get, f"set_corner_{corner}": put}


# the reference's accessor names (get_A, get_C1 .. get_T4, store_up_projectors,
# get_up_P, set_renormalized_tensors_up, get_corner_ul, ...), generated from
# the tables above
for _name in _SITE_TENSORS:
    setattr(CTMEnvironment, f"get_{_name}", _site_getter(_name))
for _d in _MOVE_TARGETS:
    for _attr, _fn in _projector_accessors(_d).items():
        _fn.__name__ = _fn.__qualname__ = _attr
        setattr(CTMEnvironment, _attr, _fn)
for _c in _ENLARGED:
    for _attr, _fn in _enlarged_accessors(_c).items():
        _fn.__name__ = _fn.__qualname__ = _attr
        setattr(CTMEnvironment, _attr, _fn)
