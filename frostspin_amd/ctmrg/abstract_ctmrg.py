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


class AbstractCTMRG:
    """
    Corner Transfer Matrix Renormalization Group for bilayer (thermal or pure)
    tensor networks on the square lattice with arbitrary unit cells.

    Leg ordering of site tensors: (physical, ancilla, up, right, down, left);
    environment tensors named clockwise from the upper left corner:

        C1-T1-T1-C2
        |  |  |   |
        T4-a--a--T2
        |  |  |   |
        C4-T3-T3-C3
    """

    def __init__(self, env, chi_target, *, verbosity=0, **ctm_params):
        self.verbosity = verbosity
        if self.verbosity > 0:
            print(f"initalize CTMRG with verbosity = {self.verbosity}")
        if type(self) is AbstractCTMRG:
            raise NotImplementedError("Cannot instantiate AbstractCTMRG")
        self._env = env
        self._site_coords = env.site_coords
        self.chi_target = int(chi_target)
        self.block_chi_ratio = float(ctm_params.get("block_chi_ratio", 1.1))
        self.ncv_ratio = float(ctm_params.get("ncv_ratio", 2.0))
        self.cutoff = float(ctm_params.get("cutoff", 1e-11))
        self.degen_ratio = float(ctm_params.get("degen_ratio", 0.9999))
        if self.verbosity > 0:
            print(self)
        if self.chi_target < 2:
            raise ValueError("chi must be larger than 2")
        if self.block_chi_ratio < 1.0:
            raise ValueError("block_chi_ratio must be larger than 1.0")
        if self.ncv_ratio < 2.0:
            raise ValueError("ncv_ratio must be larger than 2.0")
        if not (0.0 <= self.cutoff < 1.0):
            raise ValueError("cutoff must me obey 0.0 <= self.cutoff < 1.0")
        if not (0.0 < self.degen_ratio <= 1.0):
            raise ValueError("degen_ratio must obey 0.0 <= self.degen_ratio <= 1.0")

    @classmethod
    def from_elementary_tensors(cls, tiling, tensors, chi_target, *, verbosity=0, **ctm_params):
        """Start from a dummy environment built from the unit-cell site tensors."""
        if verbosity > 0:
            print("Start CTMRG from dummy environment")
        env = CTMEnvironment.from_elementary_tensors(tiling, tensors)
        return cls(env, chi_target, verbosity=verbosity, **ctm_params)

    def set_tensors(self, tensors):
        self._env.set_tensors(tensors)

    @classmethod
    def load_from_file(cls, filename, *, verbosity=0):
        ctm_params = {}
        with np.load(filename) as fin:
            chi_target = int(fin["_CTM_chi_target"])
            for k in ("block_chi_ratio", "ncv_ratio", "cutoff", "degen_ratio"):
                ctm_params[k] = float(fin[f"_CTM_{k}"])
        env = CTMEnvironment.load_from_file(filename)
        return cls(env, chi_target, verbosity=verbosity, **ctm_params)

    def save_to_file(self, filename, **additional_data):
        data = {
            "_CTM_chi_target": self.chi_target,
            "_CTM_block_chi_ratio": self.block_chi_ratio,
            "_CTM_ncv_ratio": self.ncv_ratio,
            "_CTM_cutoff": self.cutoff,
            "_CTM_degen_ratio": self.degen_ratio,
        }
        data |= self._env.get_data_dic()
        np.savez_compressed(filename, **data, **additional_data)

    @property
    def Dmax(self):
        return self._env.Dmax

    @property
    def Dmin(self):
        return self._env.Dmin

    @property
    def Lx(self):
        return self._env.Lx

    @property
    def Ly(self):
        return self._env.Ly

    @property
    def cell(self):
        return self._env.cell

    @property
    def chi_values(self):
        return self._env.chi_values

    @property
    def elementary_tensors(self):
        return self._env.elementary_tensors

    @property
    def n_sites(self):
        return self._env.n_sites

    @property
    def site_coords(self):
        return self._site_coords

    @property
    def tiling(self):
        return "\n".join("".join(s) for s in self.cell)

    def __repr__(self):
        return f"{self.symmetry()} symmetric CTMRG with Dmax = {self.Dmax} and chi_target = {self.chi_target}"

    def __str__(self):
        return "\n".join(
            (
                repr(self),
                f"chi values = {self.chi_values}",
                f"block_chi_ratio = {self.block_chi_ratio}",
                f"ncv_ratio = {self.ncv_ratio}",
                f"cutoff = {self.cutoff}",
                f"degen_ratio = {self.degen_ratio}",
                f"unit cell =\n{self._env.cell}",
            )
        )

    def symmetry(self):
        return self._env.symmetry()

    def get_A(self, x, y):
        return self._env.get_A(x, y)

    def get_C1(self, x, y):
        return self._env.get_C1(x, y)

    def get_C2(self, x, y):
        return self._env.get_C2(x, y)

    def get_C3(self, x, y):
        return self._env.get_C3(x, y)

    def get_C4(self, x, y):
        return self._env.get_C4(x, y)

    def get_T1(self, x, y):
        return self._env.get_T1(x, y)

    def get_T2(self, x, y):
        return self._env.get_T2(x, y)

    def get_T3(self, x, y):
        return self._env.get_T3(x, y)

    def get_T4(self, x, y):
        return self._env.get_T4(x, y)

    def get_corner_representations(self):
        return self._env.get_corner_representations()

    def set_symmetry(self, symmetry):
        if self.symmetry() != symmetry:
            self._env.set_symmetry(symmetry)

    # observables ---------------------------------------------------------
    def compute_PEPS_norm_log(self):
        """Log of the network norm per site (abstract_ctmrg.py:337-387)."""
        e = self._env
        logZ = 0.0
        for x, y in self._site_coords:
            nc = contract_C1234(e.get_C1(x, y), e.get_C2(x + 1, y), e.get_C4(x, y + 1), e.get_C3(x + 1, y + 1))
            nt13 = contract_T1T3(
                e.get_C1(x, y), e.get_T1(x + 1, y), e.get_C2(x + 2, y), e.get_C4(x, y + 1),
                e.get_T3(x + 1, y + 1), e.get_C3(x + 2, y + 1),
            )
            nt24 = contract_T2T4(
                e.get_C1(x, y), e.get_C2(x + 1, y), e.get_T4(x, y + 1), e.get_T2(x + 1, y + 1),
                e.get_C4(x, y + 2), e.get_C3(x + 1, y + 2),
            )
            n1234 = contract_norm(
                e.get_C1(x, y), e.get_T1(x + 1, y), e.get_C2(x + 2, y), e.get_T4(x, y + 1),
                e.get_A(x + 1, y + 1), e.get_T2(x + 2, y + 1), e.get_C4(x, y + 2),
                e.get_T3(x + 1, y + 2), e.get_C3(x + 2, y + 2),
            )
            if nc * nt13 * nt24 * n1234 < 0:
                print("Warning: negative PEPS norm")
                return np.nan
            logZ += np.log(abs(n1234)) + np.log(abs(nc)) - np.log(abs(nt13)) - np.log(abs(nt24))
        return logZ / self.n_sites

    def compute_rdm1x1(self, x, y):
        e = self._env
        return rdm.rdm_1x1(
            e.get_C1(x, y), e.get_T1(x + 1, y), e.get_C2(x + 2, y), e.get_T4(x, y + 1),
            e.get_A(x + 1, y + 1), e.get_T2(x + 2, y + 1), e.get_C4(x, y + 2),
            e.get_T3(x + 1, y + 2), e.get_C3(x + 2, y + 2),
        )

    def compute_rdm1x2(self, x, y):
        e = self._env
        return rdm.rdm_1x2(
            e.get_C1(x, y), e.get_T1(x + 1, y), e.get_T1(x + 2, y), e.get_C2(x + 3, y),
            e.get_T4(x, y + 1), e.get_A(x + 1, y + 1), e.get_A(x + 2, y + 1), e.get_T2(x + 3, y + 1),
            e.get_C4(x, y + 2), e.get_T3(x + 1, y + 2), e.get_T3(x + 2, y + 2), e.get_C3(x + 3, y + 2),
        )

    def compute_rdm2x1(self, x, y):
        e = self._env
        return rdm.rdm_2x1(
            e.get_C1(x, y), e.get_T1(x + 1, y), e.get_C2(x + 2, y), e.get_T4(x, y + 1),
            e.get_A(x + 1, y + 1), e.get_T2(x + 2, y + 1), e.get_T4(x, y + 2), e.get_A(x + 1, y + 2),
            e.get_T2(x + 2, y + 2), e.get_C4(x, y + 3), e.get_T3(x + 1, y + 3), e.get_C3(x + 2, y + 3),
        )

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
        """(abstract_ctmrg.py:321-335)"""
        e = self._env
        print("tensor shapes for C1 T1 C2 // T4 A T2 // C4 T3 C4:")
        for x, y in self._site_coords:
            print(
                f"({x},{y}):",
                e.get_C1(x, y).shape, e.get_T1(x + 1, y).shape, e.get_C2(x + 2, y).shape,
                e.get_T4(x, y + 1).shape, e.get_A(x + 1, y + 1).shape, e.get_T2(x + 2, y + 1).shape,
                e.get_C4(x, y + 3).shape, e.get_T3(x + 1, y + 2).shape, e.get_C3(x + 2, y + 2).shape,
            )

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
