"""Oracle: numpy restatement of every device entry point (TEST INFRASTRUCTURE
ONLY).

Each function reproduces the documented semantics of one kernel behind
frostspin_amd._ops (grouped GEMM, planned gathers, structural transforms,
batched SVD/QR/Cholesky, reductions) on torch CPU tensors with numpy/LAPACK.
Installed in place of the device functions it runs the product's host logic
(block bookkeeping, plan tables, descriptors) on the CPU:
* CPU tests (fixture `emulated_device`) check that host logic without a GPU;
* bench.py's cpu_baseline leg times one step of the same algorithm on the
  host cores.
The product never imports this module.
"""

import numpy as np
import scipy.linalg
import torch

from frostspin_amd import _ops


def _np(t):
    return t.numpy()


def _mixed_digits(lens, vol):
    """(vol, len(lens)) digits of 0..vol-1 in the mixed radix lens, fastest first."""
    f = np.arange(vol, dtype=np.int64)
    out = np.empty((vol, len(lens)), dtype=np.float64)
    for k, ln in enumerate(lens):
        out[:, k] = f % ln
        f //= ln
    return out


def _permute_execute(plan, src, dst):
    """Box tiles of the plan (fs_permute_plan_tiles layout): element f of a
    tile (f in source order) is read at src_base + sum_k r_k sstr_k, r_k the
    source-order digits of f, and staged at sum_k r_k lstr_k; staged element g
    is written at dst_base + sum_k r'_k dstr_k, r'_k the destination-order
    digits of g. Tiles of the same digit shape are done together: their
    offsets are one (exact, < 2^53) float64 product digits @ strides. Every
    address is checked against the buffers first."""
    K = _ops._lib.FS_PERMUTE_MAX_DIGITS
    tiles, _ = plan.tiles()
    s = _np(src) if src is not None and src.numel() else np.zeros((0,))
    d = _np(dst)
    if not tiles.shape[0]:
        return
    slen = tiles[:, 5 : 5 + 3 * K : 3]
    sstr = tiles[:, 6 : 6 + 3 * K : 3].astype(np.float64)
    lstr = tiles[:, 7 : 7 + 3 * K : 3].astype(np.float64)
    dlen = tiles[:, 5 + 3 * K : 5 + 5 * K : 2]
    dstr = tiles[:, 6 + 3 * K : 6 + 5 * K : 2].astype(np.float64)
    shapes, group = np.unique(np.concatenate([slen, dlen], axis=1), axis=0, return_inverse=True)
    group = group.reshape(-1)
    for g in range(shapes.shape[0]):
        ids_all = np.flatnonzero(group == g)
        ls = [int(x) for x in shapes[g, :K] if x > 0]
        ld = [int(x) for x in shapes[g, K:] if x > 0]
        vol = int(np.prod(ls)) if ls else 1
        assert vol == (int(np.prod(ld)) if ld else 1) == int(tiles[ids_all[0], 2])
        ns, nd = len(ls), len(ld)
        Ds, Dd = _mixed_digits(ls, vol), _mixed_digits(ld, vol)
        for c0 in range(0, ids_all.size, 2048):
            ids = ids_all[c0 : c0 + 2048]
            so = np.rint(Ds @ sstr[ids, :ns].T).astype(np.int64)  # (vol, T)
            lo = np.rint(Ds @ lstr[ids, :ns].T).astype(np.int64)
            do = np.rint(Dd @ dstr[ids, :nd].T).astype(np.int64)
            assert lo.min() >= 0 and lo.max() < vol
            live = tiles[ids, 0] >= 0
            vals = np.zeros((vol, ids.size))
            if live.any():
                addr = tiles[ids[live], 0][None, :] + so[:, live]
                assert addr.min() >= 0 and addr.max() < s.size, "tile reads outside the source"
                vals[:, live] = s[addr]
            staged = np.empty_like(vals)
            staged[lo, np.arange(ids.size)[None, :]] = vals
            daddr = tiles[ids, 1][None, :] + do
            assert daddr.min() >= 0 and daddr.max() < d.size, "tile writes outside the destination"
            d[daddr] = staged


def _view(buf, off, rows, cols, rs, cs):
    # the same bounds the kernel relies on: every addressed element is in the buffer
    last = int(off) + (int(rows) - 1) * int(rs) + (int(cols) - 1) * int(cs)
    assert int(off) >= 0 and min(int(rs), int(cs)) >= 0 and last < buf.size, "out of bounds"
    return np.lib.stride_tricks.as_strided(
        buf[int(off) :], shape=(int(rows), int(cols)), strides=(int(rs) * 8, int(cs) * 8)
    )


def grouped_gemm(groups, A, B, C, *, alpha=1.0, beta=0.0, split=False):
    n_g = len(groups)
    As, Bs, Cs = (list(x) if isinstance(x, (list, tuple)) else [x] * n_g for x in (A, B, C))
    assert len(As) == len(Bs) == len(Cs) == n_g
    for (m, n, k, ao, ars, acs, bo, brs, bcs, co, cld), A1, B1, C1 in zip(groups, As, Bs, Cs, strict=True):
        if m == 0 or n == 0:
            continue
        a, b, c = _np(A1), _np(B1), _np(C1)
        am = _view(a, ao, m, k, ars, acs)
        bm = _view(b, bo, k, n, brs, bcs)
        cm = _view(c, co, m, n, cld, 1)
        res = alpha * (am @ bm)
        cm[...] = res if beta == 0.0 else res + beta * cm


def batched_svd(descs, A, U, S, VT, work, *, max_sweeps=60):
    a, u, s, vt = _np(A), _np(U), _np(S), _np(VT)
    for m, n, ao, ars, acs, uo, so, vo, _wo in descs:
        q = min(m, n)
        am = _view(a, ao, m, n, ars, acs).copy()
        if not np.isfinite(am).all():
            # the Jacobi kernel runs its bounded sweeps on non-finite input and
            # returns non-finite factors; callers detect and redo such matrices
            u[uo : uo + m * q] = np.nan
            s[so : so + q] = np.nan
            vt[vo : vo + q * n] = np.nan
            continue
        uu, ss, vv = np.linalg.svd(am, full_matrices=False)
        u[uo : uo + m * q] = uu.reshape(-1)
        s[so : so + q] = ss
        vt[vo : vo + q * n] = vv.reshape(-1)


def batched_qr(descs, A, Q, R, work):
    a, qb = _np(A), _np(Q)
    rb = None if R is None else _np(R)
    for m, n, ao, ars, acs, qo, ro, _wo in descs:
        q = min(m, n)
        qq, rr = np.linalg.qr(_view(a, ao, m, n, ars, acs).copy(), mode="reduced")
        qb[qo : qo + m * q] = qq.reshape(-1)
        if rb is not None:
            rb[ro : ro + q * n] = rr.reshape(-1)


def _desc_rows(descs, names):
    """Descriptor tuples from either form the product passes: a sequence of
    tuples, or a dict of per-field columns (scalars broadcast)."""
    if not isinstance(descs, dict):
        return descs
    n = max(np.size(v) for v in descs.values())
    cols = [np.broadcast_to(np.asarray(descs[k]), (n,)) for k in names]
    return [tuple(c[i].item() for c in cols) for i in range(n)]


def batched_cholesky(descs, G, R, status):
    g, r, st = _np(G), _np(R), _np(status)
    descs = _desc_rows(descs, ("l", "g_off", "r_off", "shift_rel"))
    for b, (l, go, ro, shift_rel) in enumerate(descs):
        gm = g[go : go + l * l].reshape(l, l)
        gm = np.triu(gm) + np.triu(gm, 1).T
        gm = gm + shift_rel * np.trace(gm) * np.eye(l)
        try:
            rr = np.linalg.cholesky(gm).T
            st[b] = 0
        except np.linalg.LinAlgError:
            rr = np.zeros((l, l))
            st[b] = 1
        r[ro : ro + l * l] = rr.reshape(-1)


def batched_cholinv(descs, G, R, Rinv, status):
    g, st = _np(G), _np(status)
    ri = None if Rinv is None else _np(Rinv)
    r = None if R is None else _np(R)
    descs = _desc_rows(descs, ("l", "g_off", "g_stride", "n_part", "r_off", "rinv_off", "shift_rel"))
    for b, (l, go, gs, npart, ro, rio, shift_rel) in enumerate(descs):
        assert 1 <= l <= _ops.CHOLINV_MAX_ORDER
        gm = sum(g[go + p * gs : go + p * gs + l * l].reshape(l, l) for p in range(npart))
        gm = np.triu(gm) + np.triu(gm, 1).T
        gm = gm + shift_rel * np.trace(gm) * np.eye(l)
        try:
            rr = np.linalg.cholesky(gm).T
            st[b] = 0
        except np.linalg.LinAlgError:
            # the kernel stops at the failed pivot and leaves R / Rinv unwritten;
            # NaN stands for that unspecified content
            st[b] = 1
            if ri is not None:
                ri[rio : rio + l * l] = np.nan
            if r is not None:
                r[ro : ro + l * l] = np.nan
            continue
        if r is not None:
            r[ro : ro + l * l] = rr.reshape(-1)
        if ri is not None:
            ri[rio : rio + l * l] = np.linalg.inv(rr).reshape(-1)


def batched_triinv(descs, R, X, status=None):
    """X = R^-1 per upper triangular R (NaN where the status is nonzero)."""
    r, x = _np(R), _np(X)
    st = None if status is None else _np(status)
    descs = _desc_rows(descs, ("l", "r_off", "x_off"))
    for b, (l, ro, xo) in enumerate(descs):
        if st is not None and st[b] != 0:
            x[xo : xo + l * l] = np.nan
            continue
        rr = np.triu(r[ro : ro + l * l].reshape(l, l))
        with np.errstate(all="ignore"):
            x[xo : xo + l * l] = scipy.linalg.solve_triangular(
                rr, np.eye(l), lower=False, check_finite=False).reshape(-1)


def cholqr(descs, Y, Q, Rt, status, *, passes=3, shifted=True):
    """Multi-pass shifted CholeskyQR per matrix: G = Y^T Y, R = chol(G + s tr(G) I)
    (s = 11 (m l + l (l + 1)) eps on the first pass when shifted), Y <- Y R^-1;
    Q = the last Y, Rt = R_{passes-1} ... R_0; a failed factorisation sets that
    pass's status and turns the matrix into NaN."""
    y, q, st = _np(Y), _np(Q), _np(status)
    rt = None if Rt is None else _np(Rt)
    rows = _desc_rows(descs, ("m", "l", "y_off", "y_rs", "y_cs", "q_off", "rt_off"))
    n = len(rows)
    eps = np.finfo(np.float64).eps
    for i, (m, l, yo, yrs, ycs, qo, rto) in enumerate(rows):
        cur = np.array(_view(y, yo, m, l, yrs, ycs))
        rtot = None
        for p in range(passes):
            gm = cur.T @ cur
            gm = np.triu(gm) + np.triu(gm, 1).T
            shift = 11.0 * (m * l + l * (l + 1.0)) * eps if shifted and p == 0 else 0.0
            gm = gm + shift * np.trace(gm) * np.eye(l)
            try:
                r = np.linalg.cholesky(gm).T
                st[p * n + i] = 0
                x = scipy.linalg.solve_triangular(r, np.eye(l), lower=False, check_finite=False)
            except np.linalg.LinAlgError:
                st[p * n + i] = 1
                r = np.full((l, l), np.nan)
                x = r
            cur = cur @ x
            rtot = r if rtot is None else r @ rtot
        q[qo : qo + m * l] = cur.reshape(-1)
        if rt is not None:
            rt[rto : rto + l * l] = rtot.reshape(-1)


def batched_trsm_right_upper(descs, Y, R, O):
    y, r, o = _np(Y), _np(R), _np(O)
    for m, l, yo, yrs, ycs, ro, oo in descs:
        if m == 0:
            continue
        ym = _view(y, yo, m, l, yrs, ycs)
        rm = r[ro : ro + l * l].reshape(l, l)
        with np.errstate(all="ignore"):  # like the kernel: a zero pivot gives inf/nan
            x = np.zeros((m, l))
            for j in range(l):
                x[:, j] = (ym[:, j] - x[:, :j] @ rm[:j, j]) / rm[j, j]
        o[oo : oo + m * l] = x.reshape(-1)


def weighted_norm2(x, seg_off, weights):
    xs = _np(x)
    tot = 0.0
    for i, w in enumerate(weights):
        seg = xs[int(seg_off[i]) : int(seg_off[i + 1])]
        tot += w * float(seg @ seg)
    return torch.tensor([tot], dtype=torch.float64)


def scale_(x, alpha):
    x.mul_(alpha)


def scale_rsqrt_dev_(x, d):
    x.mul_(1.0 / float(np.sqrt(float(d[0]))))


def axpby_(alpha, x, beta, y):
    if beta == 0.0:
        y.copy_(alpha * x)
    else:
        y.copy_(alpha * x + beta * y)


def copy2d(descs, src, dst, *, row_w=None, col_w=None, col_idx=None):
    s, d = _np(src), _np(dst)
    rw = None if row_w is None else _np(row_w)
    cw = None if col_w is None else _np(col_w)
    ci = None if col_idx is None else _np(col_idx)
    for rows, cols, so, srs, scs, do, drs, dcs, rwo, cwo, cio, alpha, beta in descs:
        if rows == 0 or cols == 0:
            continue
        i = np.arange(rows)[:, None]
        j = np.arange(cols)[None, :]
        jj = j if ci is None else ci[int(cio) + j]
        v = alpha * s[so + i * srs + jj * scs]
        if rw is not None:
            v = v * rw[rwo + i]
        if cw is not None:
            v = v * cw[cwo + j]
        di = do + i * drs + j * dcs
        d[di] = v if beta == 0.0 else v + beta * d[di]


def gather2d(descs, idx, w, src, dst):
    from frostspin_amd import _lib

    idx = np.ascontiguousarray(idx, dtype=np.int64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    arr = (_lib.Gather2dDesc * max(len(descs), 1))(
        *[_lib.Gather2dDesc(*(int(x) for x in d[:14]), float(d[14])) for d in descs])
    _lib.check(_lib.lib.fs_gather2d_check(len(descs), arr, _lib.i64p(idx), idx.size, w.size,
                                          src.numel(), dst.numel()))
    s, d = _np(src), _np(dst)
    for rows, cols, _sr, _sc, so, srs, scs, do, drs, dcs, rio, rwo, cio, cwo, beta in descs:
        if rows == 0 or cols == 0:
            continue
        ri = idx[rio : rio + 2 * rows].reshape(rows, 2)
        rw = w[rwo : rwo + 2 * rows].reshape(rows, 2)
        ci = idx[cio : cio + 2 * cols].reshape(cols, 2)
        cw = w[cwo : cwo + 2 * cols].reshape(cols, 2)
        v = np.zeros((rows, cols))
        for a in range(2):
            for b in range(2):
                v += rw[:, a, None] * cw[None, :, b] * s[so + ri[:, a, None] * srs + ci[None, :, b] * scs]
        di = do + np.arange(rows)[:, None] * drs + np.arange(cols)[None, :] * dcs
        d[di] = v if beta == 0.0 else v + beta * d[di]


def weighted_trace(x, offs, ns, lds, weights):
    xs = _np(x)
    tot = 0.0
    for o, n, ld, w in zip(offs, ns, lds, weights, strict=True):
        tot += w * xs[int(o) + np.arange(int(n)) * (int(ld) + 1)].sum()
    return torch.tensor([tot], dtype=torch.float64)


def fill_normal_(x, seed):
    rng = np.random.default_rng(int(seed))
    x.copy_(torch.from_numpy(rng.standard_normal(x.numel())))


class EmulatedLiePlan:
    """fs_lie_plan semantics in numpy; tables are validated by the native
    host-only checker (fs_lie_plan_check) exactly as fs_lie_plan_create does."""

    def __init__(self, tasks, rows, digits, w, src_size, dst_size):
        from frostspin_amd import _lib

        self.tasks = np.ascontiguousarray(tasks, dtype=np.int64).reshape(-1, 8)
        self.rows = np.ascontiguousarray(rows, dtype=np.int64).reshape(-1, 3)
        self.digits = np.ascontiguousarray(digits, dtype=np.int64).reshape(-1, 5)
        self.w = np.ascontiguousarray(w, dtype=np.float64).reshape(-1)
        self.src_size, self.dst_size = int(src_size), int(dst_size)
        _lib.check(_lib.lib.fs_lie_plan_check(
            self.tasks.shape[0], _lib.i64p(self.tasks), self.rows.shape[0], _lib.i64p(self.rows),
            self.digits.shape[0], _lib.i64p(self.digits), self.w.size, self.src_size, self.dst_size))
        t = self.tasks
        self.flops = float((2 * t[:, 2] * t[:, 3] * t[:, 0]).sum())
        self.coeffs = float(((t[:, 2] + t[:, 3]) * t[:, 0]).sum())

    def execute(self, src, dst):
        s = _np(src) if src is not None else np.zeros((0,))
        d = _np(dst)
        for E, w_off, n_in, n_out, pad, row0, ndig, dig0 in self.tasks:
            if n_out == 0:
                continue
            e = np.arange(E)
            er_i = np.zeros_like(e)
            ec_i = np.zeros_like(e)
            er_o = np.zeros_like(e)
            ec_o = np.zeros_like(e)
            rem = e.copy()
            for radix, in_r, in_c, out_r, out_c in self.digits[dig0 + ndig - 1 : dig0 - 1 if dig0 else None : -1]:
                r = rem % radix
                rem //= radix
                er_i += r * in_r
                ec_i += r * in_c
                er_o += r * out_r
                ec_o += r * out_c
            X = np.zeros((n_in, E))
            for i in range(n_in):
                base, rs, cs = self.rows[row0 + i]
                if base >= 0:
                    idx = base + er_i * rs + ec_i * cs
                    assert idx.min() >= 0 and idx.max() < self.src_size
                    X[i] = s[idx]
            Wt = self.w[w_off : w_off + n_in * pad].reshape(n_in, pad)[:, :n_out]
            Y = Wt.T @ X
            for o in range(n_out):
                base, rs, cs = self.rows[row0 + n_in + o]
                idx = base + er_o * rs + ec_o * cs
                assert idx.min() >= 0 and idx.max() < self.dst_size
                d[idx] = Y[o]


def lie_plan_build(pattern, degen, in_block, block_off, block_rs, block_cs, weights, src_size):
    """fs_lie_plan_build in emulation: the native host tables, then the numpy plan."""
    w = weights.w
    n_in = np.ctypeslib.as_array(pattern.n_in, shape=(pattern.ntasks,))
    n_out = np.ctypeslib.as_array(pattern.n_out, shape=(pattern.ntasks,))
    tasks, rows, digits, shapes = _ops.lie_plan_tables(
        pattern, degen, in_block, block_off, block_rs, block_cs, nrows=int(n_in.sum() + n_out.sum()))
    dst = int((shapes[:, 0] * shapes[:, 1]).sum())
    return EmulatedLiePlan(tasks, rows, digits, w, src_size, dst), shapes


def _execute(self, src, dst):
    _permute_execute(self, src, dst)


class _Patcher:
    """Minimal monkeypatch stand-in that can undo its changes."""

    def __init__(self):
        self._saved = []

    def setattr(self, obj, name, value):
        self._saved.append((obj, name, getattr(obj, name)))
        setattr(obj, name, value)

    def undo(self):
        for obj, name, value in reversed(self._saved):
            setattr(obj, name, value)
        self._saved.clear()


def install_global():
    """Install the emulation process-wide; returns a callable that restores the
    device functions."""
    p = _Patcher()
    install(p)
    return p.undo


def install(monkeypatch):
    """Route frostspin_amd._ops device calls to the numpy emulation (CPU tensors)."""
    monkeypatch.setattr(_ops, "device", lambda: torch.device("cpu"))
    monkeypatch.setattr(_ops.PermutePlan, "execute", _execute)
    monkeypatch.setattr(_ops, "LiePlan", EmulatedLiePlan)
    for name in (
        "grouped_gemm", "batched_svd", "batched_qr", "batched_cholesky", "batched_cholinv",
        "batched_triinv", "cholqr", "batched_trsm_right_upper", "lie_plan_build",
        "weighted_norm2", "scale_", "scale_rsqrt_dev_",
        "axpby_", "copy2d", "gather2d", "weighted_trace", "fill_normal_",
    ):
        monkeypatch.setattr(_ops, name, globals()[name])
