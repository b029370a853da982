"""Simultaneous CTMRG schedule (ctmrg/simultaneous_ctmrg.py:13-233).

One iterate() builds the four enlarged corners of every inequivalent site,
then the projectors of the four directions from those same corners, then
renormalizes every C and T of the cell at once and replaces the environment.
The corner orders passed to construct_projectors and the coordinate offsets
of every stored projector and renormalized tensor follow the reference so
that environments agree leg by leg.
"""

from .abstract_ctmrg import AbstractCTMRG
from .ctm_renormalize import (
    renormalize_quadrant,
    renormalize_T1,
    renormalize_T2,
    renormalize_T3,
    renormalize_T4,
)


class SimultaneousCTMRG(AbstractCTMRG):
    def __repr__(self):
        return (
            f"{self.symmetry()} symmetric SimultaneousCTMRG with Dmax = {self.Dmax}"
            f" and chi_target = {self.chi_target}"
        )

    def iterate(self):
        self.construct_enlarged_corners()
        self.compute_projectors()
        self.renormalize_tensors()

    def construct_enlarged_corners(self):
        """All four enlarged corners of every site, kept in the environment
        (simultaneous_ctmrg.py:71-80)."""
        for x, y in self._site_coords:
            self.construct_enlarged_dr(x, y)
            self.construct_enlarged_ur(x, y)
            self.construct_enlarged_ul(x, y)
            self.construct_enlarged_dl(x, y)

    def compute_projectors(self):
        """Projectors of the four directions from the same corners
        (simultaneous_ctmrg.py:82-150): each direction cuts the 2x2 block of
        enlarged corners along a different bond. All sites and directions are
        independent and share their device launches."""
        e = self._env
        problems, stores = [], []
        for x, y in self._site_coords:
            dl = self.construct_enlarged_dl(x, y)
            dr = self.construct_enlarged_dr(x, y)
            ur = self.construct_enlarged_ur(x, y)
            ul = self.construct_enlarged_ul(x, y)
            problems += [
                (dr, ur, ul, dl, e.get_C2(x + 2, y + 1)),
                (dl, dr, ur, ul, e.get_C3(x + 2, y + 2).transpose()),
                (ul, dl, dr, ur, e.get_C4(x + 1, y + 2)),
                (ur, ul, dl, dr, e.get_C1(x + 1, y + 1)),
            ]
            stores += [
                (e.store_up_projectors, x + 2, y),
                (e.store_right_projectors, x + 3, y + 2),
                (e.store_down_projectors, x + 1, y + 3),
                (e.store_left_projectors, x, y + 1),
            ]
        for (store, px, py), (P, Pt) in zip(stores, self._projectors_batch(problems), strict=True):
            store(px, py, P, Pt)

    def renormalize_tensors(self):
        """New C1..C4 (projected enlarged corners) and T1..T4 of every site
        (simultaneous_ctmrg.py:152-233), then a fresh environment."""
        e = self._env
        new = [[] for _ in range(9)]
        for x, y in self._site_coords:
            A = e.get_A(x, y)
            tensors = (
                A,
                renormalize_quadrant(
                    e.get_up_P(x + 1, y - 1), self.construct_enlarged_ul(x - 1, y - 1), e.get_left_Pt(x - 1, y)
                ),
                renormalize_quadrant(
                    e.get_right_P(x + 1, y + 1), self.construct_enlarged_ur(x - 2, y - 1), e.get_up_Pt(x, y - 1)
                ),
                renormalize_quadrant(
                    e.get_down_P(x - 1, y + 1), self.construct_enlarged_dr(x - 2, y - 2), e.get_right_Pt(x + 1, y)
                ).transpose(),
                renormalize_quadrant(
                    e.get_left_P(x - 1, y - 1), self.construct_enlarged_dl(x - 1, y - 2), e.get_down_Pt(x, y + 1)
                ),
                renormalize_T1(e.get_up_Pt(x, y - 1), e.get_T1(x, y - 1), A, e.get_up_P(x + 1, y - 1)),
                renormalize_T2(e.get_right_Pt(x + 1, y), e.get_T2(x + 1, y), A, e.get_right_P(x + 1, y + 1)),
                renormalize_T3(e.get_down_Pt(x, y + 1), e.get_T3(x, y + 1), A, e.get_down_P(x - 1, y + 1)),
                renormalize_T4(e.get_left_Pt(x - 1, y), e.get_T4(x - 1, y), A, e.get_left_P(x - 1, y - 1)),
            )
            for lst, t in zip(new, tensors, strict=True):
                lst.append(t)
        self._env = type(e)(self.cell, *new)
