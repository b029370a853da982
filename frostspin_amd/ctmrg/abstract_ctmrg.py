"""CTMRG base class (ctmrg/abstract_ctmrg.py:17-668).

Same constructor, parameters, properties, I/O and observables API as the
reference so that scripts switch by changing the import. Tensors are
device-resident SymmetricTensors; every contraction is a device permute or a
grouped MFMA GEMM, every projector a batched device SVD.
"""

import numpy as np

from . import observables, rdm
from .ctm_contract import (
    contract_C1234,
    contract_dl,
    contract_dr,
    contract_norm,
    contract_T1T3,
    contract_T2T4,
    contract_ul,
    contract_ur,
)
from .ctm_environment import CTMEnvironment
from .ctm_renormalize import construct_projectors_batch


# CTMRG parameters: default and validity test (abstract_ctmrg.py:17-120); they
# are saved with the environment as _CTM_<name>
_PARAMS = {
    "block_chi_ratio": (1.1, lambda v: v >= 1.0, "block_chi_ratio must be at least 1.0"),
    "ncv_ratio": (2.0, lambda v: v >= 2.0, "ncv_ratio must be at least 2.0"),
    "cutoff": (1e-11, lambda v: 0.0 <= v < 1.0, "cutoff must lie in [0, 1)"),
    "degen_ratio": (0.9999, lambda v: 0.0 < v <= 1.0, "degen_ratio must lie in (0, 1]"),
}
# read-only views of the environment
_ENV_PROPERTIES = ("Dmax", "Dmin", "Lx", "Ly", "cell", "chi_values", "elementary_tensors", "n_sites")
_ENV_GETTERS = ("get_A", "get_C1", "get_C2", "get_C3", "get_C4", "get_T1", "get_T2", "get_T3", "get_T4",
                "get_corner_representations")


class AbstractCTMRG:
    """
    Corner Transfer Matrix Renormalization Group for bilayer (thermal or pure)
    tensor networks on the square lattice with arbitrary unit cells.

    Site tensors have legs (physical, ancilla, up, right, down, left). Around a
    window of sites the environment reads, clockwise from the upper left corner,

        C1 T1 .. T1 C2
        T4 A  .. A  T2
        ..          ..
        C4 T3 .. T3 C3

    Every rdm, norm and transfer-matrix contraction takes the tensors of such a
    window in row-major order (_window).
    """

    def __init__(self, env, chi_target, *, verbosity=0, **ctm_params):
        if type(self) is AbstractCTMRG:
            raise NotImplementedError("AbstractCTMRG is abstract: use a CTMRG schedule class")
        self.verbosity = verbosity
        self._env = env
        self._site_coords = env.site_coords
        self.chi_target = int(chi_target)
        if self.chi_target < 2:
            raise ValueError("chi_target must be at least 2")
        for name, (default, valid, msg) in _PARAMS.items():
            value = float(ctm_params.get(name, default))
            if not valid(value):
                raise ValueError(msg)
            setattr(self, name, value)
        if self.verbosity > 0:
            print(f"CTMRG ready (verbosity {self.verbosity})\n{self}")

    @classmethod
    def from_elementary_tensors(cls, tiling, tensors, chi_target, *, verbosity=0, **ctm_params):
        """Start from the trivial environment of the unit-cell site tensors."""
        env = CTMEnvironment.from_elementary_tensors(tiling, tensors)
        return cls(env, chi_target, verbosity=verbosity, **ctm_params)

    def set_tensors(self, tensors):
        self._env.set_tensors(tensors)

    @classmethod
    def load_from_file(cls, filename, *, verbosity=0):
        """Restart from an npz written by save_to_file (abstract_ctmrg.py:168-215 keys)."""
        with np.load(filename) as fin:
            chi_target = int(fin["_CTM_chi_target"])
            params = {name: float(fin[f"_CTM_{name}"]) for name in _PARAMS}
        return cls(CTMEnvironment.load_from_file(filename), chi_target, verbosity=verbosity, **params)

    def save_to_file(self, filename, **additional_data):
        data = {"_CTM_chi_target": self.chi_target}
        data.update({f"_CTM_{name}": getattr(self, name) for name in _PARAMS})
        data |= self._env.get_data_dic()
        np.savez_compressed(filename, **data, **additional_data)

    @property
    def site_coords(self):
        return self._site_coords

    @property
    def tiling(self):
        return "\n".join("".join(row) for row in self.cell)

    def __repr__(self):
        return f"{self.symmetry()} CTMRG, chi_target {self.chi_target}, Dmax {self.Dmax}"

    def __str__(self):
        params = ", ".join(f"{name} = {getattr(self, name)}" for name in _PARAMS)
        return f"{self!r}\nchi values {self.chi_values}; {params}\ncell:\n{self._env.cell}"

    def symmetry(self):
        return self._env.symmetry()

    def set_symmetry(self, symmetry):
        if self.symmetry() != symmetry:
            self._env.set_symmetry(symmetry)

    # windows of the environment ------------------------------------------------------
    def _window(self, x, y, w, h):
        """Tensors around the w x h sites whose upper-left site is (x+1, y+1),
        row by row: C1, w T1, C2 / (T4, w A, T2) x h / C4, w T3, C3."""
        e = self._env
        out = [e.get_C1(x, y), *(e.get_T1(x + 1 + i, y) for i in range(w)), e.get_C2(x + 1 + w, y)]
        for j in range(1, h + 1):
            out += [e.get_T4(x, y + j), *(e.get_A(x + 1 + i, y + j) for i in range(w)), e.get_T2(x + 1 + w, y + j)]
        out += [e.get_C4(x, y + 1 + h), *(e.get_T3(x + 1 + i, y + 1 + h) for i in range(w)),
                e.get_C3(x + 1 + w, y + 1 + h)]
        return out

    # observables ---------------------------------------------------------------------
    def compute_PEPS_norm_log(self):
        """Log of the network norm per site (abstract_ctmrg.py:337-387): with Z_wh
        the contraction of the w x h window, log Z = log Z_11 + log Z_00 -
        log Z_10 - log Z_01 averaged over the cell sites."""
        total = 0.0
        for x, y in self._site_coords:
            z = (contract_norm(*self._window(x, y, 1, 1)), contract_C1234(*self._window(x, y, 0, 0)),
                 contract_T1T3(*self._window(x, y, 1, 0)), contract_T2T4(*self._window(x, y, 0, 1)))
            if np.prod(z) < 0:
                print("Warning: negative PEPS norm")
                return np.nan
            total += np.log(abs(z[0])) + np.log(abs(z[1])) - np.log(abs(z[2])) - np.log(abs(z[3]))
        return total / self.n_sites

    def compute_rdm1x1(self, x, y):
        return rdm.rdm_1x1(*self._window(x, y, 1, 1))

    def compute_rdm1x2(self, x, y):
        return rdm.rdm_1x2(*self._window(x, y, 2, 1))

    def compute_rdm2x1(self, x, y):
        return rdm.rdm_2x1(*self._window(x, y, 1, 2))

    def compute_rdm_diag_dr(self, x, y, *, free_memory=False):
        """Sites (x+1, y+1) and (x+2, y+2) (abstract_ctmrg.py:438-455)."""
        e = self._env
        return rdm.rdm_diag_dr(
            e.get_C1(x, y), e.get_T1(x + 1, y), self.construct_enlarged_ur(x, y, free_memory=free_memory),
            e.get_T4(x, y + 1), e.get_A(x + 1, y + 1), self.construct_enlarged_dl(x, y, free_memory=free_memory),
            e.get_A(x + 2, y + 2), e.get_T2(x + 3, y + 2), e.get_T3(x + 2, y + 3), e.get_C3(x + 3, y + 3),
        )

    def compute_rdm_diag_ur(self, x, y, *, free_memory=False):
        """Sites (x+1, y+2) and (x+2, y+1) (abstract_ctmrg.py:457-474)."""
        e = self._env
        return rdm.rdm_diag_ur(
            self.construct_enlarged_ul(x, y, free_memory=free_memory), e.get_T1(x + 2, y), e.get_C2(x + 3, y),
            e.get_A(x + 2, y + 1), e.get_T2(x + 3, y + 1), e.get_T4(x, y + 2), e.get_A(x + 1, y + 2),
            self.construct_enlarged_dr(x, y, free_memory=free_memory), e.get_C4(x, y + 3), e.get_T3(x + 1, y + 3),
        )

    def compute_rdm1x1_cell(self):
        return [self.compute_rdm1x1(x, y) for x, y in self._site_coords]

    def compute_rdm_1st_neighbor_cell(self):
        rdm1x2_cell = []
        rdm2x1_cell = []
        for x, y in self._site_coords:
            rdm1x2_cell.append(self.compute_rdm1x2(x, y))
            rdm2x1_cell.append(self.compute_rdm2x1(x, y))
        return rdm1x2_cell, rdm2x1_cell

    def compute_rdm_2nd_neighbor_cell(self, *, free_memory=True):
        """Both diagonal rdms for every cell site (abstract_ctmrg.py:500-517).
        With free_memory the cached enlarged corners are dropped first and the
        ones built here are not kept."""
        if free_memory:
            self._env.reset_constructed_corners()
        rdm_dr_cell = []
        rdm_ur_cell = []
        for x, y in self._site_coords:
            rdm_dr_cell.append(self.compute_rdm_diag_dr(x, y, free_memory=free_memory))
            rdm_ur_cell.append(self.compute_rdm_diag_ur(x, y, free_memory=free_memory))
        return rdm_dr_cell, rdm_ur_cell

    def compute_transfer_spectrum_h(self, y, nvals, *, dmax_full=200, maxiter=1000, arpack_tol=0):
        """nvals leading eigenvalues of the horizontal transfer matrix at cell row y
        (abstract_ctmrg.py:519-539)."""
        e = self._env
        T1s = [e.get_T1(-x, y) for x in range(self.Lx)]
        T3s = [e.get_T3(-x, y + 1) for x in range(self.Lx)]
        return observables.compute_mps_transfer_spectrum(
            T1s, T3s, nvals, dmax_full=dmax_full, maxiter=maxiter, arpack_tol=arpack_tol, verbosity=self.verbosity
        )

    def compute_transfer_spectrum_v(self, x, nvals, *, dmax_full=200, maxiter=1000, arpack_tol=0):
        """nvals leading eigenvalues of the vertical transfer matrix at cell column x
        (abstract_ctmrg.py:541-561)."""
        e = self._env
        T1s = [e.get_T2(x + 1, -y).permute((1,), (2, 3, 0)) for y in range(self.Ly)]
        T3s = [e.get_T4(x, -y).permute((1, 2), (3, 0)) for y in range(self.Ly)]
        return observables.compute_mps_transfer_spectrum(
            T1s, T3s, nvals, dmax_full=dmax_full, maxiter=maxiter, arpack_tol=arpack_tol, verbosity=self.verbosity
        )

    def compute_corr_length_h(self, y, *, maxiter=1000, arpack_tol=0):
        """xi = -Lx / ln|lambda_2 / lambda_1| (abstract_ctmrg.py:563-570)."""
        _, v2 = self.compute_transfer_spectrum_h(y, 2, maxiter=maxiter, arpack_tol=arpack_tol)
        return -self.Lx / np.log(np.abs(v2))

    def compute_corr_length_v(self, x, *, maxiter=1000, arpack_tol=0):
        """xi = -Ly / ln|lambda_2 / lambda_1| (abstract_ctmrg.py:572-579)."""
        _, v2 = self.compute_transfer_spectrum_v(x, 2, maxiter=maxiter, arpack_tol=arpack_tol)
        return -self.Ly / np.log(np.abs(v2))

    def print_tensor_shapes(self):
        """Shapes of the 3 x 3 window of every cell site (abstract_ctmrg.py:321-335)."""
        print("tensor shapes, C1 T1 C2 // T4 A T2 // C4 T3 C3:")
        for x, y in self._site_coords:
            print(f"({x},{y}):", *(t.shape for t in self._window(x, y, 1, 1)))

    def _projectors_batch(self, problems):
        """Projector pairs of several independent (corner1, corner2, corner3,
        corner4, last_renormalized) problems, computed in shared launches."""
        return construct_projectors_batch(
            problems, self.chi_target, block_chi_ratio=self.block_chi_ratio, ncv_ratio=self.ncv_ratio,
            rtol=self.cutoff, degen_ratio=self.degen_ratio,
        )

    # enlarged corners (cached in the environment) -------------------------
    def _enlarged(self, getter, setter, build, x, y, free_memory):
        c = getter(x, y)
        if c is None:
            c = build()
            if not free_memory:
                setter(x, y, c)
        elif free_memory:
            setter(x, y, None)
        return c

    def construct_enlarged_dr(self, x, y, *, free_memory=False):
        """Enlarged down-right corner, transposed w.r.t. clockwise order
        (abstract_ctmrg.py:581-608)."""
        e = self._env
        return self._enlarged(
            e.get_corner_dr, e.set_corner_dr,
            lambda: contract_dr(e.get_A(x + 2, y + 2), e.get_T2(x + 3, y + 2), e.get_T3(x + 2, y + 3), e.get_C3(x + 3, y + 3)),
            x, y, free_memory,
        )

    def construct_enlarged_dl(self, x, y, *, free_memory=False):
        e = self._env
        return self._enlarged(
            e.get_corner_dl, e.set_corner_dl,
            lambda: contract_dl(e.get_T4(x, y + 2), e.get_A(x + 1, y + 2), e.get_C4(x, y + 3), e.get_T3(x + 1, y + 3)),
            x, y, free_memory,
        )

    def construct_enlarged_ul(self, x, y, *, free_memory=False):
        e = self._env
        return self._enlarged(
            e.get_corner_ul, e.set_corner_ul,
            lambda: contract_ul(e.get_C1(x, y), e.get_T1(x + 1, y), e.get_T4(x, y + 1), e.get_A(x + 1, y + 1)),
            x, y, free_memory,
        )

    def construct_enlarged_ur(self, x, y, *, free_memory=False):
        e = self._env
        return self._enlarged(
            e.get_corner_ur, e.set_corner_ur,
            lambda: contract_ur(e.get_T1(x + 2, y), e.get_C2(x + 3, y), e.get_A(x + 2, y + 1), e.get_T2(x + 3, y + 1)),
            x, y, free_memory,
        )


def _env_property(name):
    return property(lambda self: getattr(self._env, name), doc=f"{name} of the environment")


def _env_getter(name):
    def get(self, *args):
        return getattr(self._env, name)(*args)

    get.__name__ = get.__qualname__ = name
    return get


for _name in _ENV_PROPERTIES:
    setattr(AbstractCTMRG, _name, _env_property(_name))
for _name in _ENV_GETTERS:
    setattr(AbstractCTMRG, _name, _env_getter(_name))
