"""Sequential CTMRG schedule (ctmrg/sequential_ctmrg.py:19-211).

One iterate() = up, right, down, left moves (on a GPU, settled iterations are
recorded once as HIP graphs and replayed: frostspin_amd.ctmrg.graphs). Each move builds the projectors
of every inequivalent site from the four enlarged corners (freeing the two
corners it renormalizes) -- the sites' projector computations are independent
and share their device launches -- then renormalizes C, T, C of every site,
then swaps the new tensors in. The move table below fixes, for each direction, which
corners feed construct_projectors, which tensors are renormalized and the
coordinate offsets, exactly as in the reference so environments agree leg by
leg.
"""

from .. import _ops
from ..config import config
from .abstract_ctmrg import AbstractCTMRG
from .ctm_renormalize import (
    renormalize_C1_left,
    renormalize_C1_up,
    renormalize_C2_right,
    renormalize_C2_up,
    renormalize_C3_down,
    renormalize_C3_right,
    renormalize_C4_down,
    renormalize_C4_left,
    renormalize_T1,
    renormalize_T2,
    renormalize_T3,
    renormalize_T4,
)
from .graphs import GraphRunner


class SequentialCTMRG(AbstractCTMRG):
    def __repr__(self):
        return (
            f"{self.symmetry()} symmetric SequentialCTMRG with Dmax = {self.Dmax}"
            f" and chi_target = {self.chi_target}"
        )

    def iterate(self):
        runner = self._graph_runner()
        if runner is None:
            self.up_move()
            self.right_move()
            self.down_move()
            self.left_move()
        else:
            runner.iterate()

    def _graph_runner(self):
        """The recorder/replayer of settled iterations (frostspin_amd.ctmrg.graphs)
        on a GPU with config["ctmrg_graphs"] (and use_graphs, a per-object
        switch, not False); None for plain eager moves (also while a
        KernelTimer times every launch on its own)."""
        runner = getattr(self, "_graphs", None)
        if (not config["ctmrg_graphs"] or not getattr(self, "use_graphs", True)
                or _ops.KernelTimer.active is not None or _ops.device().type != "cuda"):
            if runner is not None:
                runner._drop()
            return None
        if runner is None:
            runner = self._graphs = GraphRunner(self)
        return runner

    def _move_projectors(self, store, problems):
        """problems: ((x, y) of the stored projectors, corner1..4, last corner)
        per site; all sites share the projector launches."""
        pairs = self._projectors_batch([p[1:] for p in problems])
        for (xy, *_), (P, Pt) in zip(problems, pairs, strict=True):
            store(*xy, P, Pt)

    def _site_corners(self, calls):
        """The enlarged corners of every site (calls: per site, a function
        returning its four corners); sites run on separate streams."""
        return _ops.run_concurrently(calls)

    def _renormalize_sites(self, store, calls):
        """Renormalized (C, T, C) of every site (calls: per site, a function
        returning (coords, nC, nT, nC)); sites run on separate streams, the
        results are stored afterwards."""
        for xy, *tensors in _ops.run_concurrently(calls):
            store(*xy, *tensors)

    def up_move(self):
        e = self._env
        sites = self._site_coords
        corners = self._site_corners([
            lambda x=x, y=y: (
                self.construct_enlarged_dr(x, y),
                self.construct_enlarged_ur(x, y, free_memory=True),
                self.construct_enlarged_ul(x, y, free_memory=True),
                self.construct_enlarged_dl(x, y),
            )
            for x, y in sites
        ])
        self._move_projectors(e.store_up_projectors, [
            ((x + 2, y), *c, e.get_C2(x + 2, y + 1)) for (x, y), c in zip(sites, corners, strict=True)
        ])

        def site(x, y):
            P = e.get_up_P(x + 1, y)
            Pt = e.get_up_Pt(x, y)
            nC1 = renormalize_C1_up(e.get_C1(x, y), e.get_T4(x, y + 1), P)
            nT1 = renormalize_T1(Pt, e.get_T1(x, y), e.get_A(x, y + 1), P)
            nC2 = renormalize_C2_up(e.get_C2(x, y), e.get_T2(x, y + 1), Pt)
            return (x, y + 1), nC1, nT1, nC2

        self._renormalize_sites(e.store_renormalized_tensors, [lambda x=x, y=y: site(x, y) for x, y in sites])
        e.set_renormalized_tensors_up()

    def right_move(self):
        e = self._env
        sites = self._site_coords
        corners = self._site_corners([
            lambda x=x, y=y: (
                self.construct_enlarged_dl(x, y),
                self.construct_enlarged_dr(x, y, free_memory=True),
                self.construct_enlarged_ur(x, y, free_memory=True),
                self.construct_enlarged_ul(x, y),
            )
            for x, y in sites
        ])
        self._move_projectors(e.store_right_projectors, [
            ((x + 3, y + 2), *c, e.get_C3(x + 2, y + 2).transpose())
            for (x, y), c in zip(sites, corners, strict=True)
        ])

        def site(x, y):
            P = e.get_right_P(x, y + 1)
            Pt = e.get_right_Pt(x, y)
            nC2 = renormalize_C2_right(e.get_C2(x, y), e.get_T1(x - 1, y), P)
            nT2 = renormalize_T2(Pt, e.get_T2(x, y), e.get_A(x - 1, y), P)
            nC3 = renormalize_C3_right(e.get_C3(x, y), e.get_T3(x - 1, y), Pt)
            return (x - 1, y), nC2, nT2, nC3

        self._renormalize_sites(e.store_renormalized_tensors, [lambda x=x, y=y: site(x, y) for x, y in sites])
        e.set_renormalized_tensors_right()

    def down_move(self):
        e = self._env
        sites = self._site_coords
        corners = self._site_corners([
            lambda x=x, y=y: (
                self.construct_enlarged_ul(x, y),
                self.construct_enlarged_dl(x, y, free_memory=True),
                self.construct_enlarged_dr(x, y, free_memory=True),
                self.construct_enlarged_ur(x, y),
            )
            for x, y in sites
        ])
        self._move_projectors(e.store_down_projectors, [
            ((x + 1, y + 3), *c, e.get_C4(x + 1, y + 2)) for (x, y), c in zip(sites, corners, strict=True)
        ])

        def site(x, y):
            P = e.get_down_P(x - 1, y)
            Pt = e.get_down_Pt(x, y)
            nC3 = renormalize_C3_down(e.get_C3(x, y), e.get_T2(x, y - 1), P)
            nT3 = renormalize_T3(Pt, e.get_T3(x, y), e.get_A(x, y - 1), P)
            nC4 = renormalize_C4_down(e.get_C4(x, y), e.get_T4(x, y - 1), Pt)
            return (x, y - 1), nC3, nT3, nC4

        self._renormalize_sites(e.store_renormalized_tensors, [lambda x=x, y=y: site(x, y) for x, y in sites])
        e.set_renormalized_tensors_down()

    def left_move(self):
        e = self._env
        sites = self._site_coords
        corners = self._site_corners([
            lambda x=x, y=y: (
                self.construct_enlarged_ur(x, y),
                self.construct_enlarged_ul(x, y, free_memory=True),
                self.construct_enlarged_dl(x, y, free_memory=True),
                self.construct_enlarged_dr(x, y),
            )
            for x, y in sites
        ])
        self._move_projectors(e.store_left_projectors, [
            ((x, y + 1), *c, e.get_C1(x + 1, y + 1)) for (x, y), c in zip(sites, corners, strict=True)
        ])

        def site(x, y):
            P = e.get_left_P(x, y - 1)
            Pt = e.get_left_Pt(x, y)
            nC4 = renormalize_C4_left(e.get_C4(x, y), e.get_T3(x + 1, y), P)
            nT4 = renormalize_T4(Pt, e.get_T4(x, y), e.get_A(x + 1, y), P)
            nC1 = renormalize_C1_left(e.get_C1(x, y), e.get_T1(x + 1, y), Pt)
            return (x + 1, y), nC4, nT4, nC1

        self._renormalize_sites(e.store_renormalized_tensors, [lambda x=x, y=y: site(x, y) for x, y in sites])
        e.set_renormalized_tensors_left()
