"""CTMRG for an AB/BA cell whose B site is the rotated conjugate of A
(ctmrg/symmetrized_ctmrg.py:19-274).

The reflection that maps A onto B maps every enlarged corner of one site onto
an enlarged corner of the other, so each constructed corner is stored twice
(the mirror costs one device permute), only the four projectors of site (0, 0)
are computed, their mirrors serve the other site, and the renormalized
tensors of site (0, 1) are permutes of those of (0, 0). Corner mirror maps,
projector storage coordinates and the env layout follow the reference.
"""

from . import rdm
from .abstract_ctmrg import AbstractCTMRG
from .ctm_contract import contract_dl, contract_dr, contract_ul, contract_ur
from .ctm_renormalize import (
    renormalize_quadrant,
    renormalize_T1,
    renormalize_T2,
    renormalize_T3,
    renormalize_T4,
)

_CORNER_MIRROR = ((1, 0, 2), (4, 3, 5))
_PROJECTOR_MIRROR = ((1, 0, 2), (3,))


def _mirror_projectors(P, Pt):
    return P.permute(*_PROJECTOR_MIRROR), Pt.permute(*_PROJECTOR_MIRROR)


class SymmetrizedCTMRG(AbstractCTMRG):
    def __repr__(self):
        return (
            f"{self.symmetry()} symmetric SymmetrizedCTMRG with Dmax = {self.Dmax}"
            f" and chi_target = {self.chi_target}"
        )

    def _mirrored_corner(self, getter, setter, mirror_setter, build, x, y, free_memory):
        """Cached enlarged corner at (x, y) whose mirror is the corner of the
        opposite kind at (x + 1, y) (symmetrized_ctmrg.py:38-170)."""
        c = getter(x, y)
        if c is None:
            c = build()
            if not free_memory:
                setter(x, y, c)
                mirror_setter(x + 1, y, c.permute(*_CORNER_MIRROR))
        elif free_memory:
            setter(x, y, None)
            mirror_setter(x + 1, y, None)
        return c

    def construct_enlarged_dr(self, x, y, *, free_memory=False):
        """Enlarged down-right corner, transposed w.r.t. clockwise order; its
        mirror is the upper-left corner of the next site."""
        e = self._env
        return self._mirrored_corner(
            e.get_corner_dr, e.set_corner_dr, e.set_corner_ul,
            lambda: contract_dr(e.get_A(x + 2, y + 2), e.get_T2(x + 3, y + 2), e.get_T3(x + 2, y + 3), e.get_C3(x + 3, y + 3)),
            x, y, free_memory,
        )

    def construct_enlarged_dl(self, x, y, *, free_memory=False):
        e = self._env
        return self._mirrored_corner(
            e.get_corner_dl, e.set_corner_dl, e.set_corner_ur,
            lambda: contract_dl(e.get_T4(x, y + 2), e.get_A(x + 1, y + 2), e.get_C4(x, y + 3), e.get_T3(x + 1, y + 3)),
            x, y, free_memory,
        )

    def construct_enlarged_ul(self, x, y, *, free_memory=False):
        e = self._env
        return self._mirrored_corner(
            e.get_corner_ul, e.set_corner_ul, e.set_corner_dr,
            lambda: contract_ul(e.get_C1(x, y), e.get_T1(x + 1, y), e.get_T4(x, y + 1), e.get_A(x + 1, y + 1)),
            x, y, free_memory,
        )

    def construct_enlarged_ur(self, x, y, *, free_memory=False):
        e = self._env
        return self._mirrored_corner(
            e.get_corner_ur, e.set_corner_ur, e.set_corner_dl,
            lambda: contract_ur(e.get_T1(x + 2, y), e.get_C2(x + 3, y), e.get_A(x + 2, y + 1), e.get_T2(x + 3, y + 1)),
            x, y, free_memory,
        )

    def iterate(self):
        self.compute_projectors()
        self.renormalize_tensors()

    def compute_projectors(self):
        """Four projector pairs from the corners of site (0, 0), computed in
        shared launches, each stored with its mirror (symmetrized_ctmrg.py:183-224)."""
        e = self._env
        dl = self.construct_enlarged_dl(0, 0)
        dr = self.construct_enlarged_dr(0, 0)
        ur = self.construct_enlarged_ur(0, 0)
        ul = self.construct_enlarged_ul(0, 0)
        (Pu, Ptu), (Pr, Ptr), (Pd, Ptd), (Pl, Ptl) = self._projectors_batch([
            (dr, ur, ul, dl, e.get_C2(0, 1)),
            (dl, dr, ur, ul, e.get_C3(0, 0).transpose()),
            (ul, dl, dr, ur, e.get_C4(1, 0)),
            (ur, ul, dl, dr, e.get_C1(1, 1)),
        ])
        e.store_up_projectors(0, 0, Pu, Ptu)
        e.store_down_projectors(0, 1, *_mirror_projectors(Pu, Ptu))
        e.store_right_projectors(1, 0, Pr, Ptr)
        e.store_left_projectors(0, 0, *_mirror_projectors(Pr, Ptr))
        e.store_down_projectors(1, 1, Pd, Ptd)
        e.store_up_projectors(0, 1, *_mirror_projectors(Pd, Ptd))
        e.store_left_projectors(0, 1, Pl, Ptl)
        e.store_right_projectors(0, 0, *_mirror_projectors(Pl, Ptl))

    def renormalize_tensors(self):
        """Renormalized tensors of site (0, 0); those of site (0, 1) are their
        mirrors (symmetrized_ctmrg.py:226-274)."""
        e = self._env
        A = e.get_A(0, 0)
        nC1 = renormalize_quadrant(e.get_up_P(1, 1), self.construct_enlarged_ul(1, 1), e.get_left_Pt(1, 0))
        nT1 = renormalize_T1(e.get_up_Pt(0, 1), e.get_T1(0, 1), A, e.get_up_P(1, 1))
        nC2 = renormalize_quadrant(e.get_right_P(1, 1), self.construct_enlarged_ur(0, 1), e.get_up_Pt(0, 1))
        nT2 = renormalize_T2(e.get_right_Pt(1, 0), e.get_T2(1, 0), A, e.get_right_P(1, 1))
        nC3 = renormalize_quadrant(
            e.get_down_P(1, 1), self.construct_enlarged_dr(0, 0), e.get_right_Pt(1, 0)
        ).transpose()
        nT3 = renormalize_T3(e.get_down_Pt(0, 1), e.get_T3(0, 1), A, e.get_down_P(1, 1))
        nC4 = renormalize_quadrant(e.get_left_P(1, 1), self.construct_enlarged_dl(1, 0), e.get_down_Pt(0, 1))
        nT4 = renormalize_T4(e.get_left_Pt(1, 0), e.get_T4(1, 0), A, e.get_left_P(1, 1))
        self._env = type(e)(
            self.cell,
            [A, e.get_A(0, 1)],
            [nC1, nC3.transpose()],
            [nC2, nC4],
            [nC3, nC1.transpose()],
            [nC4, nC2],
            [nT1, nT3.permute((3,), (1, 0, 2))],
            [nT2, nT4.permute((3,), (0, 2, 1))],
            [nT3, nT1.permute((2, 1, 3), (0,))],
            [nT4, nT2.permute((1,), (3, 2, 0))],
        )

    def compute_rdm1x2(self, x, y):
        e = self._env
        return rdm.rdm_symmetrized_1x2(
            e.get_C1(x, y), e.get_T1(x + 1, y), e.get_T4(x, y + 1), e.get_A(x + 1, y + 1),
            e.get_C4(x, y + 2), e.get_T3(x + 1, y + 2),
        )

    def compute_rdm2x1(self, x, y):
        e = self._env
        return rdm.rdm_symmetrized_2x1(
            e.get_C1(x, y), e.get_T1(x + 1, y), e.get_C2(x + 2, y), e.get_T4(x, y + 1),
            e.get_A(x + 1, y + 1), e.get_T2(x + 2, y + 1),
        )

    def compute_rdm_2nd_neighbor_cell(self, *, free_memory=True):
        """The second site's diagonal rdms are the transposes of the first's
        (symmetrized_ctmrg.py:310-325)."""
        if free_memory:
            self._env.reset_constructed_corners()
        rdm_dr = self.compute_rdm_diag_dr(0, 0, free_memory=free_memory)
        rdm_ur = self.compute_rdm_diag_ur(0, 0, free_memory=free_memory)
        return [rdm_dr, rdm_dr.transpose()], [rdm_ur, rdm_ur.transpose()]
