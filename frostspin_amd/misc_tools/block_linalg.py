"""Batched linear algebra on lists of device matrices (one launch per step).

A `Mats` is a list of matrices living in device buffers, each described by
(buffer, offset, rows, cols, row stride, col stride). All blocks of a
symmetry-sector loop go through the same launch: one grouped GEMM, one batched
Jacobi SVD. This is the device counterpart of the per-sector Python loops of
symmetric_tensor.py:857-864 and ctm_renormalize.py:253-285.
"""

from dataclasses import dataclass

import numpy as np
import torch

from .. import _ops
from ..config import config as _debug_config


@dataclass
class Mats:
    """Matrices in device memory: matrix i is element (r, c) at
    buf[off[i] + r * rs[i] + c * cs[i]]. buf is one buffer, or a list with one
    buffer per matrix (blocks gathered from several tensors; only GEMM operands
    and outputs may be of that kind)."""

    buf: object
    off: np.ndarray
    rows: np.ndarray
    cols: np.ndarray
    rs: np.ndarray
    cs: np.ndarray

    def __len__(self):
        return len(self.off)

    def bufs(self):
        """One buffer per matrix."""
        return list(self.buf) if isinstance(self.buf, list) else [self.buf] * len(self)

    @classmethod
    def packed(cls, rows, cols, buf=None):
        """Row-major matrices packed back to back in a (new) buffer."""
        rows = np.asarray(rows, dtype=np.int64).reshape(-1)
        cols = np.asarray(cols, dtype=np.int64).reshape(-1)
        sizes = rows * cols
        off = np.zeros_like(sizes)
        if sizes.size > 1:
            off[1:] = np.cumsum(sizes[:-1])
        if buf is None:
            buf = _ops.empty(int(sizes.sum()))
        return cls(buf, off, rows, cols, cols.copy(), np.ones_like(cols))

    @classmethod
    def from_tensor_blocks(cls, st, ids):
        ids = np.asarray(ids, dtype=np.int64).reshape(-1)
        rows = st._bshapes[ids, 0].copy()
        cols = st._bshapes[ids, 1].copy()
        if st._fortran:
            rs, cs = np.ones_like(rows), rows.copy()
        else:
            rs, cs = cols.copy(), np.ones_like(cols)
        return cls(st._data, st._boffs[ids].copy(), rows, cols, rs, cs)

    @classmethod
    def concat(cls, parts):
        """All matrices of several Mats, in order (per-matrix buffers if they differ)."""
        parts = list(parts)
        cat = [np.concatenate([getattr(p, f) for p in parts]).astype(np.int64) if parts else np.zeros((0,), np.int64)
               for f in ("off", "rows", "cols", "rs", "cs")]
        single = [p.buf for p in parts if not isinstance(p.buf, list)]
        if len(single) == len(parts) and all(b is single[0] for b in single) and single:
            return cls(single[0], *cat)
        return cls([b for p in parts for b in p.bufs()], *cat)

    def T(self):
        """Transposed views (no copy)."""
        return Mats(self.buf, self.off, self.cols, self.rows, self.cs, self.rs)

    def select(self, idx):
        idx = list(idx)
        buf = [self.buf[i] for i in idx] if isinstance(self.buf, list) else self.buf
        return Mats(buf, self.off[idx], self.rows[idx], self.cols[idx], self.rs[idx], self.cs[idx])


class MatProduct:
    """Matrices A_i = F0_i @ F1_i @ ... given by their factors (Mats of equal
    length, conforming shapes) and never formed unless asked (materialize).
    gemm with a product evaluates it factor by factor starting from the other
    operand's side, so a thin operand (a range-finder panel, a block of
    singular vectors) keeps every intermediate thin: A @ X costs
    sum_f 2 rows(F_f) cols(F_f) cols(X) instead of the cost of forming A."""

    def __init__(self, factors):
        self.factors = list(factors)
        assert self.factors and all(len(f) == len(self.factors[0]) for f in self.factors)
        for f, g in zip(self.factors[:-1], self.factors[1:]):
            assert not _debug_config["debug_checks"] or (f.cols == g.rows).all()

    def __len__(self):
        return len(self.factors[0])

    @property
    def rows(self):
        return self.factors[0].rows

    @property
    def cols(self):
        return self.factors[-1].cols

    def T(self):
        return MatProduct([f.T() for f in reversed(self.factors)])

    def select(self, idx):
        idx = list(idx)
        return MatProduct([f.select(idx) for f in self.factors])

    def materialize(self, *, split=False):
        """The products as packed matrices, pairwise: (F0 F1)(F2 F3)..."""
        level = list(self.factors)
        while len(level) > 1:
            nxt = [gemm(level[i], level[i + 1], split=split) for i in range(0, len(level) - 1, 2)]
            if len(level) % 2:
                nxt.append(level[-1])
            level = nxt
        return level[0]


def _dense(A):
    return A.materialize() if isinstance(A, MatProduct) else A


def gemm(A, B, *, alpha=1.0, split=False):
    """C_i = alpha * A_i @ B_i for all i in one grouped launch (one launch per
    factor when an operand is a MatProduct); C packed row-major."""
    C = Mats.packed(A.rows, B.cols)
    gemm_into(A, B, C, alpha=alpha, split=split)
    return C


def gemm_into(A, B, C, *, alpha=1.0, split=False):
    """C_i = alpha * A_i @ B_i into given row-major outputs (C_i row stride c.rs,
    column stride 1); operands and outputs may span several buffers. A product
    operand (MatProduct, not both) is applied factor by factor from the side of
    the other operand, the last step writing C."""
    if isinstance(A, MatProduct):
        assert not isinstance(B, MatProduct)
        X = B
        for f in reversed(A.factors[1:]):
            X = gemm(f, X, split=split)
        gemm_into(A.factors[0], X, C, alpha=alpha, split=split)
        return
    if isinstance(B, MatProduct):
        X = A
        for f in B.factors[:-1]:
            X = gemm(X, f, split=split)
        gemm_into(X, B.factors[-1], C, alpha=alpha, split=split)
        return
    assert len(A) == len(B) == len(C)
    # structural checks cost as much host time as the launch itself: debug mode only
    assert not _debug_config["debug_checks"] or (
        (A.cols == B.rows).all() and (C.rows == A.rows).all() and (C.cols == B.cols).all()
        and (C.cs[C.rows * C.cols > 0] == 1).all()
    )
    groups = np.stack([A.rows, B.cols, A.cols, A.off, A.rs, A.cs, B.off, B.rs, B.cs, C.off, C.rs], axis=1)
    if any(isinstance(x.buf, list) for x in (A, B, C)):
        _ops.grouped_gemm(groups, A.bufs(), B.bufs(), C.bufs(), alpha=alpha, split=split)
    elif A.buf is B.buf or A.buf.data_ptr() == B.buf.data_ptr():
        _ops.grouped_gemm(groups, A.buf, A.buf, C.buf, alpha=alpha, split=split)
    else:
        _ops.grouped_gemm(groups, A.buf, B.buf, C.buf, alpha=alpha, split=split)


def scale_slice(A, nrows, ncols, *, row_w=None, col_w=None, w_off=None):
    """Copy the leading (nrows_i x ncols_i) corner of each A_i, optionally scaling rows
    or columns by a device weight vector at w_off[i]; output packed row-major."""
    out = Mats.packed(nrows, ncols)
    descs = []
    for i in range(len(A)):
        wo = 0 if w_off is None else int(w_off[i])
        descs.append(
            (nrows[i], ncols[i], A.off[i], A.rs[i], A.cs[i], out.off[i], out.cols[i], 1,
             wo if row_w is not None else 0, wo if col_w is not None else 0, 0, 1.0, 0.0)
        )
    _ops.copy2d(descs, A.buf, out.buf, row_w=row_w, col_w=col_w)
    return out


def svd(A, *, max_sweeps=60):
    """Thin SVD of every A_i (batched one-sided Jacobi). Returns U (packed m x q),
    S (device, packed), host singular values list, VT (packed q x n)."""
    U, S, s_off, VT, finish = svd_async(A, max_sweeps=max_sweeps)
    return U, S, s_off, finish(), VT


def svd_async(A, *, max_sweeps=60):
    """svd() without the read-back: returns U, S, s_off, VT and a callable that
    downloads the singular values (host list per matrix) when called, so the
    caller can launch other work before waiting for them."""
    q = np.minimum(A.rows, A.cols)
    if q.size and q.max() > _ops.SVD_MAX_MIN_DIM:
        msg = "dense device SVD limited to min(m, n) <= 1024; use rsvd"
        raise NotImplementedError(msg)
    A = _dense(A)
    U = Mats.packed(A.rows, q)
    VT = Mats.packed(q, A.cols)
    s_off = np.zeros_like(q)
    if q.size > 1:
        s_off[1:] = np.cumsum(q[:-1])
    w_size = q * (np.maximum(A.rows, A.cols) + q)
    w_off = np.zeros_like(w_size)
    if q.size > 1:
        w_off[1:] = np.cumsum(w_size[:-1])
    S = _ops.empty(int(q.sum()))
    work = _ops.empty(int(w_size.sum()))
    descs = [
        (A.rows[i], A.cols[i], A.off[i], A.rs[i], A.cs[i], U.off[i], s_off[i], VT.off[i], w_off[i])
        for i in range(len(A))
    ]
    _ops.batched_svd(descs, A.buf, U.buf, S, VT.buf, work, max_sweeps=max_sweeps)

    def finish(host=None):
        """Host singular values per matrix; `host` is a readback of S already
        started (download_start) and waited for."""
        if host is not None:
            s_host = host.numpy()
        else:
            s_host = _ops.download(S) if S.numel() else np.zeros((0,))
        return [s_host[s_off[i] : s_off[i] + q[i]] for i in range(len(A))]

    return U, S, s_off, VT, finish


def qr(A, *, want_r=True):
    """Economic Householder QR of every A_i (one batched launch): Q packed
    (m x q), R packed (q x n) upper triangular (or None)."""
    q = np.minimum(A.rows, A.cols)
    Q = Mats.packed(A.rows, q)
    R = Mats.packed(q, A.cols) if want_r else None
    w_size = A.rows * A.cols + q + A.rows * q
    w_off = np.zeros_like(w_size)
    if q.size > 1:
        w_off[1:] = np.cumsum(w_size[:-1])
    work = _ops.empty(int(w_size.sum()))
    descs = [
        (A.rows[i], A.cols[i], A.off[i], A.rs[i], A.cs[i], Q.off[i],
         R.off[i] if R is not None else 0, w_off[i])
        for i in range(len(A))
    ]
    _ops.batched_qr(descs, A.buf, Q.buf, None if R is None else R.buf, work)
    return Q, R


def off_parts(counts):
    """Exclusive running sum of counts (start of each matrix's parts)."""
    out = np.zeros_like(counts)
    if counts.size > 1:
        out[1:] = np.cumsum(counts[:-1])
    return out


GRAM_ROWS_PER_PART = 192
GRAM_MAX_PARTS = 16


def gram(Y):
    """Y_i^T Y_i for every i as a split-K grouped GEMM: rows of Y_i are cut into
    n_i chunks, each chunk's l x l product is a separate group (so even l x l
    outputs fill the device); partials are stored back to back (l*l apart) and
    summed by the consumer. Returns (Mats of the first partials, n_i)."""
    nparts = np.clip((Y.rows + GRAM_ROWS_PER_PART - 1) // GRAM_ROWS_PER_PART, 1, GRAM_MAX_PARTS)
    l2 = Y.cols * Y.cols
    sizes = nparts * l2
    off = np.zeros_like(sizes)
    if sizes.size > 1:
        off[1:] = np.cumsum(sizes[:-1])
    buf = _ops.empty(int(sizes.sum()))
    # part p of matrix i: rows floor(p m / n) .. floor((p + 1) m / n)
    mat = np.repeat(np.arange(len(Y)), nparts)
    part = np.arange(mat.size) - np.repeat(off_parts(nparts), nparts)
    m, npi = Y.rows[mat], nparts[mat]
    r0, r1 = part * m // npi, (part + 1) * m // npi
    a_off = Y.off[mat] + r0 * Y.rs[mat]
    l = Y.cols[mat]
    groups = np.stack([l, l, r1 - r0, a_off, Y.cs[mat], Y.rs[mat], a_off, Y.rs[mat], Y.cs[mat],
                       off[mat] + part * l2[mat], l], axis=1)
    A = Y.buf
    _ops.grouped_gemm(groups, A, A, buf)
    return Mats(buf, off, Y.cols.copy(), Y.cols.copy(), Y.cols.copy(), np.ones_like(Y.cols)), nparts


def cholqr(Y, *, shifted=True, passes=3, want_r=False, statuses=None):
    """Orthonormal basis of the column space of each Y_i (m x l, m >= l) by
    (shifted) CholeskyQR: per pass G = Y^T Y (grouped GEMM), R = chol(G + s I)
    and R^-1 (one fused LDS kernel), Y <- Y R^-1 (grouped GEMM). Shifted
    CholeskyQR3 (Fukaya et al., SIAM J. Sci. Comput. 42 (2020) A477) stays
    defined up to cond(Y) ~ 1/eps. Matrices whose Cholesky breaks down anyway
    are finished by Householder QR. With want_r, also returns the triangular
    factor R (l x l, Y = Q R) as the product of the passes' factors.

    statuses: if a list, the per-pass device status vectors are appended to it
    instead of being read back (no host synchronisation); matrices whose
    factorisation broke down are then left to the caller to redo.

    With every l_i <= 128 the passes run as one native call (fs_cholqr_f64:
    the same launches, issued without returning to Python between them),
    except under a KernelTimer, which times each launch on its own."""
    if NATIVE_CHOLQR and _ops.KernelTimer.active is None and _chain_applies(Y, passes, want_r):
        return _cholqr_chain(Y, shifted, passes, want_r, statuses)
    eps = np.finfo(np.float64).eps
    cur = Y
    Rtot = None
    for p in range(passes):
        G, nparts = gram(cur)
        status = _ops.empty(len(cur), dtype=torch.int32)  # every factorisation kernel writes it
        cols = cur.cols.astype(np.float64)
        if shifted and p == 0:
            shift = 11.0 * (cur.rows * cols + cols * (cols + 1.0)) * eps
        else:
            shift = np.zeros_like(cols)
        if len(cur) and cur.cols.max() <= _ops.CHOLINV_MAX_ORDER:
            # factor (one workgroup per matrix, four columns per barrier), then
            # the blocked inverse, then the solve as one grouped GEMM
            R = Mats.packed(cur.cols, cur.cols)
            Rinv = Mats.packed(cur.cols, cur.cols)
            _ops.batched_cholinv(
                {"l": cur.cols, "g_off": G.off, "g_stride": cur.cols * cur.cols, "n_part": nparts,
                 "r_off": R.off, "rinv_off": 0, "shift_rel": shift},
                G.buf, R.buf, None, status,
            )
            _ops.batched_triinv({"l": cur.cols, "r_off": R.off, "x_off": Rinv.off}, R.buf, Rinv.buf, status)
            out = gemm(cur, Rinv)
            if not want_r:
                R = None
        else:
            R = Mats.packed(cur.cols, cur.cols)
            if np.any(nparts > 1):
                G = gemm(cur.T(), cur)
            _ops.batched_cholesky(
                {"l": cur.cols, "g_off": G.off, "r_off": R.off, "shift_rel": shift}, G.buf, R.buf, status
            )
            out = Mats.packed(cur.rows, cur.cols)
            _ops.batched_trsm_right_upper(
                [(cur.rows[i], cur.cols[i], cur.off[i], cur.rs[i], cur.cs[i], R.off[i], out.off[i])
                 for i in range(len(cur))],
                cur.buf, R.buf, out.buf,
            )
        if want_r:
            Rtot = R if Rtot is None else gemm(R, Rtot)
        if statuses is not None:
            statuses.append(status)
            cur = out
            continue
        bad = np.nonzero(_ops.download(status))[0]
        if bad.size:
            _householder_splice(Y, bad, out, Rtot, want_r)
        cur = out
    return (cur, Rtot) if want_r else cur


NATIVE_CHOLQR = True  # tests switch it off to compare with the launch-by-launch path


def _chain_applies(Y, passes, want_r):
    return (
        len(Y) > 0
        and not isinstance(Y.buf, list)
        and 1 <= passes <= 3
        and (passes >= 2 or not want_r)
        and Y.rows.min() >= 1
        and Y.cols.min() >= 1
        and Y.cols.max() <= _ops.CHOLINV_MAX_ORDER
    )


def _cholqr_chain(Y, shifted, passes, want_r, statuses):
    n = len(Y)
    Q = Mats.packed(Y.rows, Y.cols)
    Rt = Mats.packed(Y.cols, Y.cols) if want_r else None
    status = _ops.empty(passes * n, dtype=torch.int32)
    _ops.cholqr(
        {"m": Y.rows, "l": Y.cols, "y_off": Y.off, "y_rs": Y.rs, "y_cs": Y.cs, "q_off": Q.off,
         "rt_off": Rt.off if want_r else 0},
        Y.buf, Q.buf, None if Rt is None else Rt.buf, status, passes=passes, shifted=shifted,
    )
    if statuses is not None:
        statuses.extend(status.view(passes, n).unbind(0))
    else:
        bad = np.nonzero(_ops.download(status).reshape(passes, n).any(axis=0))[0]
        if bad.size:
            _householder_splice(Y, bad, Q, Rt, want_r)
    return (Q, Rt) if want_r else Q


def _householder_splice(Y, bad, out, Rtot, want_r):
    """Breakdown (numerically rank-deficient beyond the shift): Householder QR
    of the original inputs of those matrices, copied over their results."""
    Qh, Rh = qr(Y.select(bad), want_r=want_r)
    for src, dst in ((Qh, out), (Rh, Rtot)):
        if src is None:
            continue
        descs = [
            (src.rows[j], src.cols[j], src.off[j], src.rs[j], src.cs[j], dst.off[i], dst.cols[i], 1,
             0, 0, 0, 1.0, 0.0)
            for j, i in enumerate(bad)
        ]
        _ops.copy2d(descs, src.buf, dst.buf)


def _orthonormal_columns(Y, passes=3, statuses=None, robust=False):
    """Orthonormal basis of the column space of each Y_i (Householder if robust)."""
    if robust:
        return qr(Y, want_r=False)[0]
    return cholqr(Y, passes=passes, statuses=statuses)


def rsvd(A, ks, *, oversampling=None, power_iters=2, seed=0, robust=False, certify=True,
         _oversampling_per_matrix=None):
    """Randomized truncated SVD (range finder + power iterations, Halko et al.,
    SIAM Rev. 53 (2011) Alg. 4.4/5.1) of every A_i, keeping l_i = k_i + p_i
    vectors. Returns U (m x l), S, s_off, s_list, VT (l x n) like svd().

    Orthonormalisations use shifted CholeskyQR without reading its status back
    (no host synchronisation inside the chain); the statuses are read once with
    the singular values, and any matrix whose factorisation broke down is redone
    with Householder QR (robust=True) and spliced into the results."""
    U, S, s_off, VT, finish = rsvd_async(A, ks, oversampling=oversampling, power_iters=power_iters,
                                         seed=seed, robust=robust, certify=certify,
                                         _oversampling_per_matrix=_oversampling_per_matrix)
    _ops.wait_current()
    return U, S, s_off, finish(), VT


# A leading singular triplet (s_i, u_i, v_i) of the randomized SVD is accepted
# when ||A v_i - s_i u_i|| <= RSVD_RESIDUAL_TOL * s_1 (u_i^T A = s_i v_i^T holds
# by construction, so this residual bounds the error of the pair); the
# reference's ARPACK runs with tol = 0, i.e. to machine precision.
RSVD_RESIDUAL_TOL = 1e-10
# Redo a matrix with the exact Jacobi SVD only up to this min(m, n): the
# one-workgroup Jacobi kernel's cost grows as min(m, n)^3 (about 50 ms at 256,
# tens of seconds near 1024 on slowly decaying spectra); larger matrices are
# redone by a wider randomized SVD.
EXACT_REDO_MAX = 256
# matrices certified / decomposed again, and the largest relative residual seen
RSVD_STATS = {"certified": 0, "refined": 0, "uncertified": 0, "max_rel_residual": 0.0}


def rsvd_async(A, ks, *, oversampling=None, power_iters=2, seed=0, robust=False, certify=True,
               _oversampling_per_matrix=None):
    """rsvd() with its one host readback (singular values, CholeskyQR
    statuses and residual norms) started but not waited for: returns U, S,
    s_off, VT and a callable giving the host singular values once the current
    stream has been synchronized (it splices the redone matrices, if any, into
    U, S and VT).

    certify: the k_i leading triplets of every matrix are checked against
    RSVD_RESIDUAL_TOL; a matrix that fails is decomposed again (exact Jacobi
    SVD up to the dense limit, more power iterations above) before its values
    are returned."""
    ks = np.asarray(ks, dtype=np.int64)
    mins = np.minimum(A.rows, A.cols)
    if _oversampling_per_matrix is not None:
        p = np.asarray(_oversampling_per_matrix, dtype=np.int64)
    else:
        p = np.maximum(10, ks // 2) if oversampling is None else np.full_like(ks, oversampling)
    ls = np.minimum(ks + p, mins)
    omega = Mats.packed(A.cols, ls)
    _ops.fill_normal_(omega.buf, seed)
    statuses = None if robust else []
    Y = gemm(A, omega)
    # intermediate bases only need to be well conditioned (Halko et al. sec. 4.5):
    # two CholeskyQR passes; the final basis gets the full three
    for _ in range(power_iters):
        Q = _orthonormal_columns(Y, passes=2, statuses=statuses, robust=robust)
        Z = gemm(A.T(), Q)
        Qz = _orthonormal_columns(Z, passes=2, statuses=statuses, robust=robust)
        Y = gemm(A, Qz)
    Q = _orthonormal_columns(Y, statuses=statuses, robust=robust)
    B = gemm(Q.T(), A)  # l x n
    # B^T = Q2 R2, then the small l x l SVD of R2^T = Ur S Vr^T: B = Ur S (Q2 Vr)^T
    if robust:
        Q2, R2 = qr(B.T())
    else:
        Q2, R2 = cholqr(B.T(), want_r=True, statuses=statuses)
    Ur, S, s_off, Vrt, finish_small = svd_async(R2.T())
    U = gemm(Q, Ur)  # m x l
    VT = gemm(Vrt, Q2.T())  # l x n
    s_h = _ops.download_start(S)
    st_h = _ops.download_start(torch.stack(statuses)) if statuses else None
    kc = np.minimum(ks, ls)
    res = _residual_norms2(A, U, S, s_off, VT, kc) if certify else None
    r_h = _ops.download_start(res[0]) if res is not None and res[0].numel() else None

    def finish(replay=False, redone=None):
        """Host singular values per matrix, once the stream has passed the
        readbacks. replay: the readback buffers were written again by a
        recorded graph; returns None where the eager path would decompose a
        matrix again (it never launches work then). redone: a list that
        records whether the eager path did."""
        s_list = finish_small(s_h)
        failed = np.zeros((0,), dtype=np.int64)
        if st_h is not None:
            failed = st_h.numpy().any(axis=0).nonzero()[0]
            if failed.size:
                if replay:
                    return None
                if redone is not None:
                    redone.append("breakdown")
                _splice_rsvd(A, ks, failed, U, S, s_off, s_list, VT, oversampling, power_iters, seed)
        if r_h is not None:
            r2 = r_h.numpy()
            rel = np.zeros((len(A),))
            for j in range(len(A)):
                if kc[j] and j not in failed and s_list[j][0] > 0:
                    rel[j] = np.sqrt(max(r2[res[1][j] : res[1][j] + kc[j]].max(), 0.0)) / s_list[j][0]
            bad = [j for j in range(len(A)) if rel[j] > RSVD_RESIDUAL_TOL]
            if replay and bad:
                return None
            RSVD_STATS["certified"] += int((kc > 0).sum())
            RSVD_STATS["refined"] += len(bad)
            RSVD_STATS["max_rel_residual"] = max(RSVD_STATS["max_rel_residual"], float(rel.max(initial=0.0)))
            if bad:
                if redone is not None:
                    redone.append("residual")
                _refine_rsvd(A, ks, ls, np.array(bad, dtype=np.int64), U, S, s_off, s_list, VT, power_iters, seed)
        return s_list

    return U, S, s_off, VT, finish


def _residual_norms2(A, U, S, s_off, VT, ks):
    """||A v_i - s_i u_i||^2 for the ks[j] leading triplets of every matrix j,
    on the device: (packed vector, offset of each matrix's values)."""
    r_off = off_parts(ks)
    out = _ops.empty(int(ks.sum()))
    live = [j for j in range(len(A)) if ks[j] > 0]
    if not live:
        return out, r_off
    Al = A.select(live)
    VTk = Mats(VT.buf, VT.off[live], ks[live], VT.cols[live], VT.rs[live], VT.cs[live])
    AV = gemm(Al, VTk.T())  # m x k: A v_i
    _ops.copy2d(
        [(A.rows[j], ks[j], U.off[j], U.rs[j], U.cs[j], AV.off[t], AV.rs[t], 1, 0, s_off[j], 0, -1.0, 1.0)
         for t, j in enumerate(live)],
        U.buf, AV.buf, col_w=S,
    )
    G = gemm(AV.T(), AV)  # k x k: diagonal = squared residual norms
    _ops.copy2d(
        [(ks[j], 1, G.off[t], G.cols[t] + 1, 0, r_off[j], 1, 0, 0, 0, 0, 1.0, 0.0) for t, j in enumerate(live)],
        G.buf, out,
    )
    return out, r_off


def _refine_rsvd(A, ks, ls, bad, U, S, s_off, s_list, VT, power_iters, seed):
    """Matrices whose leading triplets missed the residual tolerance are
    decomposed once more and their leading l_i triplets replace theirs: by the
    exact Jacobi SVD up to EXACT_REDO_MAX, else by a randomized SVD with
    a Krylov-sized basis (l = 2 k + 10, like ARPACK's ncv) and 3 power_iters + 2
    power iterations. The second result is kept whatever its residual (flat
    spectra converge slowly); RSVD_STATS counts those left above the
    tolerance."""
    mins = np.minimum(A.rows, A.cols)
    exact = bad[mins[bad] <= EXACT_REDO_MAX]
    wide = bad[mins[bad] > EXACT_REDO_MAX]
    parts = []
    if exact.size:
        U2, S2, s_off2, s2_list, VT2 = svd(A.select(exact))
        parts.append((exact, U2, S2, s_off2, s2_list, VT2))
    if wide.size:
        kw = ks[wide]
        over = np.maximum(kw + 10, ls[wide] - kw)
        U2, S2, s_off2, s2_list, VT2 = rsvd(A.select(wide), kw, oversampling=None, power_iters=3 * power_iters + 2,
                                            seed=seed + 1, certify=False, _oversampling_per_matrix=over)
        kc = np.minimum(kw, np.minimum(kw + over, mins[wide]))
        r, r_off = _residual_norms2(A.select(wide), U2, S2, s_off2, VT2, kc)
        r2 = _ops.download(r) if r.numel() else np.zeros((0,))
        for t in range(wide.size):
            if kc[t] and np.sqrt(max(r2[r_off[t] : r_off[t] + kc[t]].max(), 0.0)) > RSVD_RESIDUAL_TOL * s2_list[t][0]:
                RSVD_STATS["uncertified"] = RSVD_STATS.get("uncertified", 0) + 1
        parts.append((wide, U2, S2, s_off2, s2_list, VT2))
    for ids, U2, S2, s_off2, s2_list, VT2 in parts:
        l = ls[ids]
        pairs = list(enumerate(ids))
        _ops.copy2d([(U.rows[i], l[t], U2.off[t], U2.rs[t], U2.cs[t], U.off[i], U.rs[i], 1, 0, 0, 0, 1.0, 0.0)
                     for t, i in pairs], U2.buf, U.buf)
        _ops.copy2d([(l[t], VT.cols[i], VT2.off[t], VT2.rs[t], VT2.cs[t], VT.off[i], VT.rs[i], 1, 0, 0, 0, 1.0, 0.0)
                     for t, i in pairs], VT2.buf, VT.buf)
        _ops.copy2d([(l[t], 1, s_off2[t], 1, 0, s_off[i], 1, 0, 0, 0, 0, 1.0, 0.0) for t, i in pairs], S2, S)
        for t, i in pairs:
            s_list[i] = np.array(s2_list[t][: l[t]], copy=True)


def _splice_rsvd(A, ks, failed, U, S, s_off, s_list, VT, oversampling, power_iters, seed):
    """Redo the failed matrices with Householder orthonormalisation and copy the
    results over theirs (same shapes)."""
    U2, S2, s_off2, s_list2, VT2 = rsvd(
        A.select(failed), ks[failed], oversampling=oversampling, power_iters=power_iters,
        seed=seed, robust=True,
    )
    for src, dst in ((U2, U), (VT2, VT)):
        descs = [
            (src.rows[j], src.cols[j], src.off[j], src.rs[j], src.cs[j], dst.off[i], dst.cols[i], 1,
             0, 0, 0, 1.0, 0.0)
            for j, i in enumerate(failed)
        ]
        _ops.copy2d(descs, src.buf, dst.buf)
    descs = [(1, len(s_list2[j]), s_off2[j], 1, 1, s_off[i], 1, 1, 0, 0, 0, 1.0, 0.0) for j, i in enumerate(failed)]
    _ops.copy2d(descs, S2, S)
    for j, i in enumerate(failed):
        s_list[i] = s_list2[j]
