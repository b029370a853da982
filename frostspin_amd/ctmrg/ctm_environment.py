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


class CTMEnvironment:
    """Container for CTMRG environment tensors, CTMRG leg conventions."""

    def __init__(self, cell, As, C1s, C2s, C3s, C4s, T1s, T2s, T3s, T4s):
        self._cell = cell
        self._site_coords, self._indices = find_unique_coords(cell)
        self._n_sites = self._site_coords.shape[0]
        self._Ly, self._Lx = cell.shape
        self._unique_As = tuple(As)
        self._unique_C1s = list(C1s)
        self._unique_C2s = list(C2s)
        self._unique_C3s = list(C3s)
        self._unique_C4s = list(C4s)
        self._unique_T1s = list(T1s)
        self._unique_T2s = list(T2s)
        self._unique_T3s = list(T3s)
        self._unique_T4s = list(T4s)
        self._symmetry = type(As[0]).symmetry()
        self.check_consistency()
        self.reset_constructed_corners()
        self._reset_temp_lists()

    @property
    def Dmax(self):
        return max(max(A.shape[2:]) for A in self._unique_As)

    @property
    def Dmin(self):
        return min(min(A.shape[2:]) for A in self._unique_As)

    @property
    def Lx(self):
        return self._Lx

    @property
    def Ly(self):
        return self._Ly

    @property
    def cell(self):
        return self._cell

    @property
    def chi_values(self):
        s = set()
        for C in self._unique_C1s + self._unique_C3s:
            s.add(C.shape[0])
            s.add(C.shape[1])
        return sorted(s)

    @property
    def elementary_tensors(self):
        return self._unique_As

    @property
    def n_sites(self):
        return self._n_sites

    @property
    def site_coords(self):
        return self._site_coords

    def check_consistency(self):
        ST = type(self._unique_As[0])
        groups = (
            (self._unique_As, 6, "A"), (self._unique_C1s, 2, "C1"), (self._unique_C2s, 2, "C2"),
            (self._unique_C3s, 2, "C3"), (self._unique_C4s, 2, "C4"), (self._unique_T1s, 4, "T1"),
            (self._unique_T2s, 4, "T2"), (self._unique_T3s, 4, "T3"), (self._unique_T4s, 4, "T4"),
        )
        for tensors, ndim, name in groups:
            if len(tensors) != self._n_sites:
                raise ValueError(f"Invalid {name} number")
            for t in tensors:
                if type(t) is not ST:
                    raise ValueError(f"Invalid {name} type")
                if t.ndim != ndim:
                    raise ValueError(f"Invalid {name} ndim")
        for x, y in self._site_coords:
            A = self.get_A(x, y)
            Ay = self.get_A(x, y - 1)
            if A.signature[2] == Ay.signature[4]:
                raise ValueError(f"signature mismatch between {(x, y)} and {(x, y - 1)}")
            if (np.asarray(A.col_reps[0]) != np.asarray(Ay.col_reps[2])).any():
                raise ValueError(f"rep mismatch between {(x, y)} and {(x, y - 1)}")
            Ax = self.get_A(x + 1, y)
            if A.signature[3] == Ax.signature[5]:
                raise ValueError(f"signature mismatch between {(x + 1, y)} and {(x, y)}")
            if (np.asarray(A.col_reps[1]) != np.asarray(Ax.col_reps[3])).any():
                raise ValueError(f"rep mismatch between {(x + 1, y)} and {(x, y)}")

    def get_corner_representations(self):
        seen, unique = set(), []
        for C in self._unique_C1s + self._unique_C3s:
            for r in (C.row_reps[0], C.col_reps[0]):
                key = tuple(np.asarray(r).ravel())
                if key not in seen:
                    seen.add(key)
                    unique.append(r)
        ST = type(self._unique_As[0])
        return sorted(unique, key=lambda r: (ST.representation_dimension(r), tuple(np.asarray(r).ravel())))

    @classmethod
    def from_elementary_tensors(cls, tiling, tensors):
        cell = parse_tiling(tiling)
        if len(tensors) != len(set(cell.flat)):
            raise ValueError("Incompatible cell and tensor number")
        env = [initialize_dummy_env(A) for A in tensors]
        C1s, T1s, C2s, T2s, C3s, T3s, C4s, T4s = (list(x) for x in zip(*env, strict=True))
        return cls(cell, tensors, C1s, C2s, C3s, C4s, T1s, T2s, T3s, T4s)

    @classmethod
    def load_from_file(cls, savefile):
        """npz schema of ctm_environment.py:239-281."""
        names = ("A", "C1", "C2", "C3", "C4", "T1", "T2", "T3", "T4")
        lists = {k: [] for k in names}
        with np.load(savefile) as data:
            cell = data["_CTM_cell"]
            ST = get_symmetric_tensor_type(str(data["_CTM_symmetry"][()]))
            for i in range(len(set(cell.flat))):
                for k in names:
                    lists[k].append(ST.load_from_dic(data, prefix=f"_CTM_{k}_{i}"))
        return cls(cell, *(lists[k] for k in names))

    def symmetry(self):
        return self._symmetry

    def get_data_dic(self):
        data = {"_CTM_cell": self._cell, "_CTM_symmetry": self._symmetry}
        for i in range(self._n_sites):
            for k, lst in (
                ("A", self._unique_As), ("C1", self._unique_C1s), ("C2", self._unique_C2s),
                ("C3", self._unique_C3s), ("C4", self._unique_C4s), ("T1", self._unique_T1s),
                ("T2", self._unique_T2s), ("T3", self._unique_T3s), ("T4", self._unique_T4s),
            ):
                data |= lst[i].get_data_dic(prefix=f"_CTM_{k}_{i}")
        return data

    def reset_constructed_corners(self):
        self._corners_ul = [None] * self._n_sites
        self._corners_ur = [None] * self._n_sites
        self._corners_dl = [None] * self._n_sites
        self._corners_dr = [None] * self._n_sites

    def set_tensors(self, tensors):
        if len(tensors) != self._n_sites:
            raise ValueError("Incompatible cell and tensors")
        for i, A in enumerate(tensors):
            if not A.match_representations(self._unique_As[i]):
                raise ValueError("Incompatible representation for new tensor")
        self._unique_As = tuple(tensors)
        forget_site_permutes()
        self.reset_constructed_corners()

    def set_symmetry(self, symmetry):
        if symmetry != self._symmetry:
            self.reset_constructed_corners()
            self._symmetry = symmetry
            self._unique_As = tuple(A.cast(symmetry) for A in self._unique_As)
            for name in ("C1", "C2", "C3", "C4", "T1", "T2", "T3", "T4"):
                attr = f"_unique_{name}s"
                setattr(self, attr, [t.cast(symmetry) for t in getattr(self, attr)])

    def _unique_index(self, x, y):
        return self._indices[x % self._Lx, y % self._Ly]

    def get_tensor_type(self, x, y):
        return self._cell[x % self._Lx, y % self._Ly]

    def get_A(self, x, y):
        return self._unique_As[self._unique_index(x, y)]

    def get_C1(self, x, y):
        return self._unique_C1s[self._unique_index(x, y)]

    def get_T1(self, x, y):
        return self._unique_T1s[self._unique_index(x, y)]

    def get_C2(self, x, y):
        return self._unique_C2s[self._unique_index(x, y)]

    def get_T2(self, x, y):
        return self._unique_T2s[self._unique_index(x, y)]

    def get_C3(self, x, y):
        return self._unique_C3s[self._unique_index(x, y)]

    def get_T3(self, x, y):
        return self._unique_T3s[self._unique_index(x, y)]

    def get_C4(self, x, y):
        return self._unique_C4s[self._unique_index(x, y)]

    def get_T4(self, x, y):
        return self._unique_T4s[self._unique_index(x, y)]

    # projectors of the current move
    def _store(self, d, x, y, P, Pt):
        j = self._unique_index(x, y)
        self._proj[d][0][j] = P
        self._proj[d][1][j] = Pt

    def store_up_projectors(self, x, y, P, Pt):
        self._store("up", x, y, P, Pt)

    def store_right_projectors(self, x, y, P, Pt):
        self._store("right", x, y, P, Pt)

    def store_down_projectors(self, x, y, P, Pt):
        self._store("down", x, y, P, Pt)

    def store_left_projectors(self, x, y, P, Pt):
        self._store("left", x, y, P, Pt)

    def get_up_P(self, x, y):
        return self._proj["up"][0][self._unique_index(x, y)]

    def get_up_Pt(self, x, y):
        return self._proj["up"][1][self._unique_index(x, y)]

    def get_right_P(self, x, y):
        return self._proj["right"][0][self._unique_index(x, y)]

    def get_right_Pt(self, x, y):
        return self._proj["right"][1][self._unique_index(x, y)]

    def get_down_P(self, x, y):
        return self._proj["down"][0][self._unique_index(x, y)]

    def get_down_Pt(self, x, y):
        return self._proj["down"][1][self._unique_index(x, y)]

    def get_left_P(self, x, y):
        return self._proj["left"][0][self._unique_index(x, y)]

    def get_left_Pt(self, x, y):
        return self._proj["left"][1][self._unique_index(x, y)]

    # cached enlarged corners
    def get_corner_ul(self, x, y):
        return self._corners_ul[self._unique_index(x, y)]

    def get_corner_ur(self, x, y):
        return self._corners_ur[self._unique_index(x, y)]

    def get_corner_dl(self, x, y):
        return self._corners_dl[self._unique_index(x, y)]

    def get_corner_dr(self, x, y):
        return self._corners_dr[self._unique_index(x, y)]

    def set_corner_ul(self, x, y, ul):
        self._corners_ul[self._unique_index(x, y)] = ul

    def set_corner_ur(self, x, y, ur):
        self._corners_ur[self._unique_index(x, y)] = ur

    def set_corner_dl(self, x, y, dl):
        self._corners_dl[self._unique_index(x, y)] = dl

    def set_corner_dr(self, x, y, dr):
        self._corners_dr[self._unique_index(x, y)] = dr

    # renormalized tensors of the current move
    def store_renormalized_tensors(self, x, y, nCX, nT, nCY):
        j = self._unique_index(x, y)
        self._nCX[j] = nCX
        self._nT[j] = nT
        self._nCY[j] = nCY

    def _reset_temp_lists(self):
        n = self._n_sites
        self._proj = {d: ([None] * n, [None] * n) for d in ("up", "right", "down", "left")}
        self._nCX = [None] * n
        self._nT = [None] * n
        self._nCY = [None] * n

    def set_renormalized_tensors_up(self):
        self._unique_C1s, self._unique_T1s, self._unique_C2s = self._nCX, self._nT, self._nCY
        self._reset_temp_lists()

    def set_renormalized_tensors_right(self):
        self._unique_C2s, self._unique_T2s, self._unique_C3s = self._nCX, self._nT, self._nCY
        self._reset_temp_lists()

    def set_renormalized_tensors_down(self):
        self._unique_C3s, self._unique_T3s, self._unique_C4s = self._nCX, self._nT, self._nCY
        self._reset_temp_lists()

    def set_renormalized_tensors_left(self):
        self._unique_C4s, self._unique_T4s, self._unique_C1s = self._nCX, self._nT, self._nCY
        self._reset_temp_lists()
