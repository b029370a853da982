"""CTMRG base class (ctmrg/abstract_ctmrg.py:17-668).

Same constructor, parameters, properties, I/O and observables API as the
reference so that scripts switch by changing the import. Tensors are
device-resident SymmetricTensors; every contraction is a device permute or a
grouped MFMA GEMM, every projector a batched device SVD.
"""

import numpy as np

from . import rdm
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

    def compute_rdm1x1_cell(self):
        return [self.compute_rdm1x1(x, y) for x, y in self._site_coords]

    def compute_rdm_1st_neighbor_cell(self):
        rdm1x2_cell = []
        rdm2x1_cell = []
        for x, y in self._site_coords:
            rdm1x2_cell.append(self.compute_rdm1x2(x, y))
            rdm2x1_cell.append(self.compute_rdm2x1(x, y))
        return rdm1x2_cell, rdm2x1_cell

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
