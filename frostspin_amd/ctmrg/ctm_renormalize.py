"""CTMRG projectors and renormalisation (ctmrg/ctm_renormalize.py:6-341).

construct_projectors runs every symmetry sector in the same launches:
R = c1 @ c2, Rt = c3 @ c4 and M = R @ Rt either formed as three grouped FP64
MFMA GEMMs over all shared sectors or, when a cost model says it is cheaper
(large sectors), kept as the factor chain c1 c2 c3 c4 and applied factor by
factor to the thin operands; then one batched SVD of all M blocks (exact
Jacobi for blocks up to 1024, randomized range finder with power iterations
above, where the reference calls ARPACK), then the multiplet-aware cut on the
host, then P = s^-1/2 U^T R and Pt = Rt V s^-1/2 as grouped GEMMs with the
s^-1/2 scaling fused into the slicing copy.
"""

import numpy as np
import torch

from .. import _ops
from ..misc_tools import block_linalg as bl
from ..misc_tools.svd_tools import find_block_cuts
from .ctm_contract import site_permute

# Blocks up to SMALL_DENSE (smallest side) always use the exact Jacobi SVD; above
# DENSE_SVD_LIMIT the randomized truncated SVD is the only option.
SMALL_DENSE = 96
DENSE_SVD_LIMIT = 1024


def construct_projectors(
    corner1,
    corner2,
    corner3,
    corner4,
    chi_target,
    last_renormalized,
    *,
    block_chi_ratio=1.1,
    ncv_ratio=2.0,
    atol=0.0,
    rtol=None,
    degen_ratio=1.0,
    rsvd_seed=0,
):
    """Isometries P (mid -> corner2 columns) and Pt keeping about chi_target
    singular values of R @ Rt; interface of ctm_renormalize.py:6-298."""
    return construct_projectors_batch(
        [(corner1, corner2, corner3, corner4, last_renormalized)], chi_target,
        block_chi_ratio=block_chi_ratio, ncv_ratio=ncv_ratio, atol=atol, rtol=rtol,
        degen_ratio=degen_ratio, rsvd_seed=rsvd_seed,
    )[0]


class _CutDecision:
    """The host step of a projector batch between its one wait and the
    projector launches: singular values read back, redone decompositions,
    cuts. The launches after it depend on the cuts only (the s^-1/2 weights
    are gathered on the device), so a CTMRG graph recorder may replay them
    before this decision is re-evaluated: replay() then tells whether the
    cuts equal the recorded ones."""

    def __init__(self):
        self.evaluate = None
        self.cuts = None
        self.non_empty = None
        self.redone = []

    def replay(self):
        """After a replay of the recorded launches: True when the re-evaluated
        cuts equal the recorded ones, False when the eager path would decide
        otherwise (other cuts, or a decomposition to redo)."""
        res = self.evaluate(True)
        return res is not None and np.array_equal(res[0], self.cuts)


def _inv_sqrt_kept(S_of, cuts, non_empty):
    """s^-1/2 of the kept singular values of every non-empty sector, packed in
    that order, computed on the device: S_of[i] = (device buffer, offset) of
    sector i's values; one gather from those buffers, then 1 / sqrt."""
    if not len(non_empty):
        return _ops.empty(0)
    srcs, pos = [], []
    for i in non_empty:
        t, off = S_of[i]
        k = next((q for q, x in enumerate(srcs) if x is t), None)
        if k is None:
            srcs.append(t)
            k = len(srcs) - 1
        pos.append((k, off, int(cuts[i])))
    base = np.zeros((len(srcs),), dtype=np.int64)
    if len(srcs) > 1:
        base[1:] = np.cumsum([x.numel() for x in srcs[:-1]])
    idx = np.concatenate([base[k] + off + np.arange(c, dtype=np.int64) for k, off, c in pos])
    S_all = srcs[0] if len(srcs) == 1 else torch.cat(srcs)
    idx_dev = _ops.upload_async(idx, dtype=torch.int64, structural=True)  # fixed by the cuts
    return 1.0 / torch.sqrt(torch.index_select(S_all, 0, idx_dev))


def construct_projectors_batch(
    problems,
    chi_target,
    *,
    block_chi_ratio=1.1,
    ncv_ratio=2.0,
    atol=0.0,
    rtol=None,
    degen_ratio=1.0,
    rsvd_seed=0,
    diagnostics=None,
):
    """construct_projectors for several independent (corner1, corner2, corner3,
    corner4, last_renormalized) problems at once: the sectors of all problems
    share every launch (operands gathered from the different corners through the
    per-group operand bases of the grouped GEMM), while the truncation is decided
    per problem exactly as for a single call. Returns a list of (P, Pt).
    diagnostics, if a dict, receives the sectors (problem, irrep, block_chi),
    the kept counts and the singular values handed to the cut."""
    sectors = []  # (problem, shared irrep, block_chi, corner block indices)
    for pi, (corner1, corner2, corner3, corner4, last) in enumerate(problems):
        shared = sorted(
            set(corner1.block_irreps.tolist())
            & set(corner2.block_irreps.tolist())
            & set(corner3.block_irreps.tolist())
            & set(corner4.block_irreps.tolist())
        )
        shared = np.array(shared, dtype=corner2.block_irreps.dtype)
        ind = [c.block_irreps.searchsorted(shared) for c in (corner1, corner2, corner3, corner4)]
        # number of values per sector, from the last renormalized corner (:244-250)
        block_chi = np.zeros(shared.shape, dtype=int)
        last_irreps = last.block_irreps
        for bi, i in enumerate(last_irreps.searchsorted(shared)):
            if i < last_irreps.size and shared[bi] == last_irreps[i]:
                block_chi[bi] = last._bshapes[i, 1]
        block_chi = np.maximum(block_chi + 10, (block_chi_ratio * block_chi).astype(int))
        block_chi = np.minimum(chi_target + 10, block_chi)
        for bi in range(len(shared)):
            sectors.append((pi, shared[bi], int(block_chi[bi]), [int(x[bi]) for x in ind]))

    # projector construction stays FP64 whatever the contraction precision
    split = False
    # sectors are ordered by problem: gather each corner's blocks problem by problem
    cmats = [
        bl.Mats.concat(
            [
                bl.Mats.from_tensor_blocks(prob[c], [sec[3][c] for sec in sectors if sec[0] == pi])
                for pi, prob in enumerate(problems)
            ]
        )
        for c in range(4)
    ]
    n = len(sectors)
    block_chi = np.array([sec[2] for sec in sectors], dtype=int)
    mins = np.minimum(cmats[0].rows, cmats[3].cols)
    ncv = (ncv_ratio * block_chi).astype(int) + 10
    # the reference uses ARPACK (block_chi values) where min(shape) >= max(100, 4 ncv)
    sparse_ref = mins >= np.maximum(100, 4 * ncv)
    # values needed per sector: block_chi, the reference's own estimate from the
    # last renormalized corner (what it passes to ARPACK). Small blocks are
    # decomposed exactly (all values, as the reference's LAPACK branch); larger
    # ones go through the randomized range finder (GEMM-dominated) for block_chi
    # values, which is what the reference's sparse branch computes.
    k_need = np.minimum(block_chi, mins)
    l_need = np.minimum(k_need + np.maximum(10, k_need // 2), mins)
    dense = (mins <= SMALL_DENSE) | ((l_need >= 0.8 * mins) & (mins <= DENSE_SVD_LIMIT))
    dense_ids = [i for i in range(n) if dense[i]]
    rand_ids = [i for i in range(n) if not dense[i]]
    if _implicit_pays(cmats, dense_ids, rand_ids, l_need, k_need):
        # R = c1 c2, Rt = c3 c4 and M = R Rt applied factor by factor to the
        # thin operands; only the dense-branch sectors are formed
        R = bl.MatProduct(cmats[:2])
        Rt = bl.MatProduct(cmats[2:])
        M = bl.MatProduct(cmats)
    else:
        R = bl.gemm(cmats[0], cmats[1], split=split)
        Rt = bl.gemm(cmats[2], cmats[3], split=split)
        M = bl.gemm(R, Rt, split=split)
    # every singular value the truncation needs reaches the host in one wait:
    # the readbacks are started where the values are produced, and the side
    # stream joins the current stream before the wait
    dense_done = rand_done = None
    if dense_ids:
        # the exact SVDs of the small sectors are independent of the randomized
        # chain of the large ones: on a GPU they run on a side stream meanwhile
        side = _ops.side_stream() if rand_ids else None
        if side is not None:
            main = torch.cuda.current_stream()
            side.wait_stream(main)
            with torch.cuda.stream(side):
                U, S, s_off, VT, finish = bl.svd_async(M.select(dense_ids))
                s_h = _ops.download_start(S)
            for t in (U.buf, S, VT.buf):
                t.record_stream(main)
            main.wait_stream(side)
        else:
            U, S, s_off, VT, finish = bl.svd_async(M.select(dense_ids))
            s_h = _ops.download_start(S)
        dense_done = (U, VT, finish, s_h, S, s_off)
    if rand_ids:
        rand_done = bl.rsvd_async(M.select(rand_ids), k_need[rand_ids], seed=rsvd_seed)

    # The reference hands find_block_cuts every singular value of a sector on
    # its dense branch and block_chi values on its ARPACK branch (:96-103); the
    # default relative cutoff counts them all (svd_tools.py:45-48).
    ref_len = np.where(sparse_ref, np.minimum(block_chi, mins), mins)
    by_problem = [[i for i in range(n) if sectors[i][0] == pi] for pi in range(len(problems))]
    decision = _CutDecision()

    def decide(replay):
        """Singular values read back -> per-sector cuts. replay: the readback
        buffers were written again by a recorded graph; returns None where the
        eager path would launch more work (a redone decomposition)."""
        U_of, s_of, VT_of, S_of = [None] * n, [None] * n, [None] * n, [None] * n
        if rand_done is not None:
            U, S, s_off, VT, finish = rand_done
            s_list = finish(replay=replay, redone=decision.redone)
            if s_list is None:
                return None
            for j, i in enumerate(rand_ids):
                U_of[i] = (U, j)
                VT_of[i] = (VT, j)
                s_of[i] = s_list[j]
                S_of[i] = (S, int(s_off[j]))
        if dense_done is not None:
            U, VT, finish, s_h, S, s_off = dense_done
            s_list = finish(s_h)
            for j, i in enumerate(dense_ids):
                U_of[i] = (U, j)
                VT_of[i] = (VT, j)
                s_of[i] = s_list[j]
                S_of[i] = (S, int(s_off[j]))
        # the randomized SVD's oversampled tail is not accurate: only the k_need
        # leading values enter the cut (block_chi on the reference's ARPACK branch)
        for i in rand_ids:
            s_of[i] = s_of[i][: k_need[i]]

        def cut_problem(pi):
            ids = by_problem[pi]
            if not ids:
                return
            dims = [problems[pi][1].irrep_dimension(sectors[i][1]) for i in ids]
            rt = rtol if rtol is not None else float(ref_len[ids].sum()) * np.finfo(np.float64).eps
            cuts[ids] = find_block_cuts([s_of[i] for i in ids], chi_target, dims=dims, atol=atol, rtol=rt,
                                        degen_ratio=degen_ratio)

        # multiplet-aware cut, per problem
        cuts = np.zeros((n,), dtype=np.int64)
        for pi in range(len(problems)):
            cut_problem(pi)
        # A dense-branch sector served by the randomized SVD holds only its leading
        # values; while the cut takes all of them (a sector growing faster than
        # block_chi allows), the decision may depend on the values not computed:
        # such sectors are decomposed again (exact Jacobi for small sectors, a
        # randomized SVD of twice as many values above) and their problem is cut
        # again, until every such cut stops inside the computed values (the values
        # beyond the cut do not change the decision; the default cutoff already
        # counts every value of the sector).
        while True:
            short = [i for i in range(n)
                     if not sparse_ref[i] and len(s_of[i]) < mins[i] and cuts[i] >= len(s_of[i])]
            if not short:
                break
            if replay:
                return None
            decision.redone.append("short")
            exact = [i for i in short if mins[i] <= bl.EXACT_REDO_MAX]
            wider = [i for i in short if mins[i] > bl.EXACT_REDO_MAX]
            if exact:
                U, S, s_off, VT, finish = bl.svd_async(M.select(exact))
                s_list = finish()
                for j, i in enumerate(exact):
                    U_of[i], VT_of[i], s_of[i] = (U, j), (VT, j), s_list[j]
                    S_of[i] = (S, int(s_off[j]))
            if wider:
                ks = np.array([min(int(mins[i]), 2 * len(s_of[i])) for i in wider], dtype=np.int64)
                U, S, s_off, VT, finish = bl.rsvd_async(M.select(wider), ks, seed=rsvd_seed)
                _ops.wait_current()
                s_list = finish()
                for j, i in enumerate(wider):
                    U_of[i], VT_of[i], s_of[i] = (U, j), (VT, j), s_list[j][: ks[j]]
                    S_of[i] = (S, int(s_off[j]))
            for pi in sorted({sectors[i][0] for i in short}):
                cut_problem(pi)
        return cuts, s_of, U_of, VT_of, S_of

    decision.evaluate = decide
    _ops.host_wait(decision)
    cuts, s_of, U_of, VT_of, S_of = decide(False)
    decision.cuts = cuts.copy()
    if diagnostics is not None:
        diagnostics.update(
            sectors=[(sec[0], sec[1], sec[2]) for sec in sectors], cuts=cuts.copy(),
            values=[np.array(v, copy=True) for v in s_of], sparse_ref=sparse_ref.copy(), mins=mins.copy(),
        )
    non_empty = cuts.nonzero()[0]
    decision.non_empty = non_empty

    # s^-1/2 of the kept values, one device vector for all sectors, from the
    # device copies of the singular values (the launches after the wait then
    # depend on the cuts only, not on values read back)
    w_off = np.zeros((len(non_empty),), dtype=np.int64)
    if len(non_empty) > 1:
        w_off[1:] = np.cumsum(cuts[non_empty])[:-1]
    _ops.host_resume(decision)
    w_dev = _inv_sqrt_kept(S_of, cuts, non_empty)

    # Us = U[:, :cut] * s^-1/2 and Vs = s^-1/2 * VT[:cut, :] per sector (grouped by source buffer)
    kept = [int(cuts[i]) for i in non_empty]
    Us = _scaled([U_of[i] for i in non_empty], kept, w_dev, w_off, cols=True)
    Vs = _scaled([VT_of[i] for i in non_empty], kept, w_dev, w_off, cols=False)

    # P = Us^T R (cut x n2) and Pt = Rt Vs^T (m3 x cut), one buffer per problem
    ne_of = [[k for k, i in enumerate(non_empty) if sectors[i][0] == pi] for pi in range(len(problems))]
    Rk = R.select(non_empty)
    Rtk = Rt.select(non_empty)
    P_out = [bl.Mats.packed(np.array([kept[k] for k in ks], np.int64), Rk.cols[ks]) for ks in ne_of]
    Pt_out = [bl.Mats.packed(Rtk.rows[ks], np.array([kept[k] for k in ks], np.int64)) for ks in ne_of]
    order = [k for ks in ne_of for k in ks]
    if order:
        bl.gemm_into(Us.T().select(order), Rk.select(order), bl.Mats.concat(P_out), split=split)
        bl.gemm_into(Rtk.select(order), Vs.T().select(order), bl.Mats.concat(Pt_out), split=split)

    out = []
    for pi, (_, corner2, _, _, _) in enumerate(problems):
        ks = ne_of[pi]
        sec_ids = non_empty[ks] if ks else np.zeros((0,), np.int64)
        block_irreps = np.array([sectors[i][1] for i in sec_ids], dtype=corner2.block_irreps.dtype)
        mid_rep = corner2.init_representation(cuts[sec_ids], block_irreps)
        nrr2 = corner2.n_row_reps
        sp = np.zeros((nrr2 + 1,), dtype=bool)
        sp[1:] = corner2.signature[nrr2:]
        spt = np.ones((nrr2 + 1,), dtype=bool)
        spt[:-1] = ~corner2.signature[nrr2:]
        ST = type(corner2)
        Pm, Ptm = P_out[pi], Pt_out[pi]
        P = ST._build(
            (mid_rep,), corner2.col_reps, Pm.buf, np.stack([Pm.rows, Pm.cols], axis=1), Pm.off, block_irreps, sp,
        ).transpose()
        Pt = ST._build(
            corner2.col_reps, (mid_rep,), Ptm.buf, np.stack([Ptm.rows, Ptm.cols], axis=1), Ptm.off, block_irreps, spt,
        )
        out.append((P, Pt))
    return out


# Cost model of a grouped FP64 GEMM launch on the device: at least
# GEMM_LAUNCH_S (launch and tail latency of a small launch), else its flops at
# the rate measured on MI355X for its kind: GEMM_RATE_SQUARE for the corner
# products (chi=256 SU(2) sectors: 37 TFLOP/s), GEMM_RATE_THIN for products
# with a range-finder panel (l = 20..69 columns: 17 TFLOP/s).
GEMM_LAUNCH_S = 20e-6
GEMM_RATE_SQUARE = 37e12
GEMM_RATE_THIN = 17e12
# None: the cost model below picks formed or factored products per call;
# True / False forces one form (tests compare the two).
FACTORED = None


def _implicit_pays(cmats, dense_ids, rand_ids, l_need, k_need, power_iters=2):
    """Whether the projector step is cheaper with M = c1 c2 c3 c4 (and R, Rt)
    applied factor by factor than formed. Formed: three launches over every
    sector (R, Rt, M), then one launch per product with M in the randomized SVD
    (range finder, 2 per power iteration, the projection B = Q^T M and the
    residual check) and one per projector. Factored: the dense-branch sectors
    formed (three launches over them), four launches per product with M and two
    per projector. Large sectors favour the factored form (chi=256 SU(2): the
    three formed launches cost ~1.7 ms, the extra thin launches ~1.2 ms); small
    ones the formed form (each extra launch costs its latency)."""

    def launch(flop, rate=GEMM_RATE_THIN):
        return max(GEMM_LAUNCH_S, flop / rate) if flop > 0 else 0.0

    if FACTORED is not None:
        return FACTORED
    if not rand_ids:
        return False
    rows = [c.rows.astype(np.float64) for c in cmats]
    cols = [c.cols.astype(np.float64) for c in cmats]
    every = np.arange(len(rows[0]))
    dense = np.asarray(dense_ids, dtype=np.int64)
    rand = np.asarray(rand_ids, dtype=np.int64)

    def form(ids):  # R, Rt, M over the sectors ids
        return (launch(2.0 * (rows[0] * cols[0] * cols[1])[ids].sum(), GEMM_RATE_SQUARE)
                + launch(2.0 * (rows[2] * cols[2] * cols[3])[ids].sum(), GEMM_RATE_SQUARE)
                + launch(2.0 * (rows[0] * cols[1] * cols[3])[ids].sum(), GEMM_RATE_SQUARE))

    n_apply = 1 + 2 * power_iters + 2
    width = l_need.astype(np.float64)[rand]
    thin_m = 2.0 * (rows[0] * cols[3])[rand] * width
    thin_f = [2.0 * (rows[f] * cols[f])[rand] * width for f in range(4)]
    kept = k_need.astype(np.float64)
    proj_m = [2.0 * (kept * rows[0] * cols[1]).sum(), 2.0 * (kept * rows[2] * cols[3]).sum()]
    proj_f = [2.0 * (kept * rows[f] * cols[f]).sum() for f in range(4)]
    formed = form(every) + n_apply * launch(thin_m.sum()) + sum(launch(x) for x in proj_m)
    factored = ((form(dense) if dense.size else 0.0) + n_apply * sum(launch(x.sum()) for x in thin_f)
                + sum(launch(x) for x in proj_f))
    return factored < formed


def _scaled(parts, cuts, w_dev, w_off, *, cols):
    """Leading cut columns of U (cols=True) or rows of VT, scaled by s^-1/2, packed."""
    if not parts:
        return bl.Mats.packed(np.zeros((0,), np.int64), np.zeros((0,), np.int64))
    rows_out, cols_out, descs = [], [], []
    for k, ((mats, j), c) in enumerate(zip(parts, cuts, strict=True)):
        if cols:
            rows_out.append(int(mats.rows[j]))
            cols_out.append(c)
        else:
            rows_out.append(c)
            cols_out.append(int(mats.cols[j]))
    out = bl.Mats.packed(rows_out, cols_out)
    # group descriptors by source buffer (U/VT of dense and randomized paths)
    by_buf = {}
    for k, ((mats, j), c) in enumerate(zip(parts, cuts, strict=True)):
        by_buf.setdefault(id(mats.buf), (mats.buf, []))[1].append((k, mats, j))
    for buf, items in by_buf.values():
        descs = []
        for k, mats, j in items:
            descs.append(
                (rows_out[k], cols_out[k], mats.off[j], mats.rs[j], mats.cs[j], out.off[k],
                 out.cols[k], 1, 0 if cols else w_off[k], w_off[k] if cols else 0, 0, 1.0, 0.0)
            )
        if cols:
            _ops.copy2d(descs, buf, out.buf, col_w=w_dev)
        else:
            _ops.copy2d(descs, buf, out.buf, row_w=w_dev)
    return out


# conventions: renormalize_C(C, T, P), renormalize_T(Pt, T, A, P)
def renormalize_T(Pt, T, A, P):
    """Absorb one site into an edge and project it back (ctm_renormalize.py:307-328);
    cost 2 chi^2 D^4 (d a D^2 + chi) dense-equivalent flop."""
    nT = T @ P.permute((2,), (0, 1, 3))
    nT = A @ nT.permute((0, 3), (1, 4, 2, 5))
    nT = nT.permute((0, 1, 4, 5), (2, 3, 6, 7))
    nT = site_permute(A, (0, 1, 4, 5), (2, 3)).dagger() @ nT
    nT = Pt.transpose() @ nT.permute((3, 1, 4), (2, 0, 5))
    return nT.normalize_()


def renormalize_corner_P(C, T, P):
    """(ctm_renormalize.py:331-353), 2 chi^3 D^2 flop."""
    nC = (T @ C).permute((0,), (1, 2, 3)) @ P
    return nC.normalize_()


def renormalize_corner_Pt(C, T, Pt):
    """(ctm_renormalize.py:356-378)"""
    nC = (T @ C).permute((0,), (1, 2, 3)) @ Pt
    return nC.normalize_()


def renormalize_quadrant(P, quadrant, Pt):
    nc = P.transpose() @ quadrant @ Pt
    return nc.normalize_()


def renormalize_C1_up(C1, T4, P):
    return renormalize_corner_P(C1.transpose(), T4.permute((3, 1, 2), (0,)), P).transpose()


def renormalize_T1(Pt, T1, A, P):
    nT1 = renormalize_T(Pt, T1.permute((1, 2, 3), (0,)), site_permute(A, (0, 1, 4, 5), (2, 3)), P)
    return nT1.permute((3,), (1, 2, 0))


def renormalize_C2_up(C2, T2, Pt):
    return renormalize_corner_Pt(C2, T2.permute((1, 2, 3), (0,)), Pt)


def renormalize_C2_right(C2, T1, P):
    return renormalize_corner_P(C2.transpose(), T1.permute((3, 1, 2), (0,)), P).transpose()


def renormalize_T2(Pt, T2, A, P):
    nT2 = renormalize_T(Pt, T2.permute((2, 3, 0), (1,)), site_permute(A, (0, 1, 5, 2), (3, 4)), P)
    return nT2.permute((0,), (3, 1, 2))


def renormalize_C3_right(C3, T3, Pt):
    return renormalize_corner_Pt(C3.transpose(), T3.permute((3, 0, 1), (2,)), Pt).transpose()


def renormalize_C3_down(C3, T2, P):
    return renormalize_corner_P(C3, T2.permute((0, 2, 3), (1,)), P)


def renormalize_T3(Pt, T3, A, P):
    nT3 = renormalize_T(Pt, T3.permute((0, 1, 2), (3,)), site_permute(A, (0, 1, 2, 3), (4, 5)), P)
    return nT3.permute((1, 2, 0), (3,))


def renormalize_C4_down(C4, T4, Pt):
    return renormalize_corner_Pt(C4, T4.permute((0, 1, 2), (3,)), Pt)


def renormalize_C4_left(C4, T3, P):
    return renormalize_corner_P(C4.transpose(), T3.permute((2, 0, 1), (3,)), P).transpose()


def renormalize_T4(Pt, T4, A, P):
    nT4 = renormalize_T(Pt, T4.permute((1, 2, 3), (0,)), site_permute(A, (0, 1, 3, 4), (5, 2)), P)
    return nT4.permute((3,), (1, 2, 0))


def renormalize_C1_left(C1, T1, Pt):
    return renormalize_corner_Pt(C1, T1.permute((0, 1, 2), (3,)), Pt)
