"""Batched device linear algebra (grouped GEMM, Householder QR, Jacobi SVD,
randomized truncated SVD) on packed block lists, against numpy/scipy."""

import numpy as np
import pytest
import scipy.linalg as lg

from frostspin_amd import _ops
from frostspin_amd.misc_tools import block_linalg as bl

BACKENDS = [
    pytest.param("emulated", id="emulated"),
    pytest.param("gpu", id="gpu", marks=pytest.mark.gpu),
]


@pytest.fixture(params=BACKENDS)
def backend(request):
    if request.param == "emulated":
        request.getfixturevalue("emulated_device")
    else:
        request.getfixturevalue("gpu")
    return request.param


def _pack(mats):
    rows = [m.shape[0] for m in mats]
    cols = [m.shape[1] for m in mats]
    buf = _ops.upload(np.concatenate([m.reshape(-1) for m in mats]))
    return bl.Mats.packed(rows, cols, buf)


def _unpack(M):
    h = _ops.download(M.buf)
    out = []
    for i in range(len(M)):
        idx = M.off[i] + np.arange(M.rows[i])[:, None] * M.rs[i] + np.arange(M.cols[i])[None, :] * M.cs[i]
        out.append(h[idx])
    return out


def _decaying(rng, m, n, decay):
    u, _ = np.linalg.qr(rng.standard_normal((m, min(m, n))))
    v, _ = np.linalg.qr(rng.standard_normal((n, min(m, n))))
    s = decay ** np.arange(min(m, n))
    return (u * s) @ v.T, s


def test_qr_matches_lapack(backend):
    rng = np.random.default_rng(0)
    mats = [rng.standard_normal(sh) for sh in [(7, 3), (3, 7), (40, 40), (130, 20), (1, 1)]]
    mats[2][:, 3] = 0.0
    Q, R = bl.qr(_pack(mats))
    for a, q, r in zip(mats, _unpack(Q), _unpack(R), strict=True):
        qr_, rr_ = lg.qr(a, mode="economic")
        assert np.abs(q @ r - a).max() < 1e-13 * max(1, np.abs(a).max())
        assert np.abs(np.tril(r, -1)).max() == 0.0
        assert np.abs(q.T @ q - np.eye(q.shape[1])).max() < 1e-13
        # LAPACK sign convention: same factors up to rounding
        assert np.abs(np.abs(np.diag(r)) - np.abs(np.diag(rr_))).max() < 1e-12


def test_rsvd_top_values(backend):
    rng = np.random.default_rng(1)
    mats, specs = [], []
    for m, n in [(300, 260), (150, 400), (257, 257)]:
        a, s = _decaying(rng, m, n, 0.8)
        mats.append(a)
        specs.append(s)
    A = _pack(mats)
    ks = [20, 15, 30]
    U, S, s_off, s_list, VT = bl.rsvd(A, ks, power_iters=2, seed=3)
    us, vts = _unpack(U), _unpack(VT)
    for i, a in enumerate(mats):
        k = ks[i]
        s = s_list[i]
        assert np.abs(s[:k] - specs[i][:k]).max() < 1e-10
        # leading singular vectors reproduce the rank-k truncation
        ak = (us[i][:, :k] * s[:k]) @ vts[i][:k]
        ref = (np.linalg.svd(a)[0][:, :k] * specs[i][:k]) @ np.linalg.svd(a)[2][:k]
        # subspace error ~ (s_{l+1} / s_k)^(2q+1) s_k (Halko et al. Cor. 10.10)
        l_i = min(k + max(10, k // 2), min(a.shape))
        bound = (specs[i][l_i] / specs[i][k - 1]) ** 5 * specs[i][k - 1]
        assert np.abs(ak - ref).max() < 10 * bound


def test_dense_svd_lds_and_global_paths(backend):
    rng = np.random.default_rng(2)
    mats = [rng.standard_normal(sh) for sh in [(12, 9), (90, 80), (400, 150)]]
    U, S, s_off, s_list, VT = bl.svd(_pack(mats))
    for a, u, s, vt in zip(mats, _unpack(U), s_list, _unpack(VT), strict=True):
        ref = np.linalg.svd(a, compute_uv=False)
        assert np.abs(s - ref).max() < 1e-13 * ref[0]
        assert np.abs((u * s) @ vt - a).max() < 1e-12 * ref[0]


def test_cholqr_q_and_r(backend):
    """Shifted CholeskyQR3: fused Cholesky+inverse path (l <= 128), the
    Cholesky+substitution path (l > 128), ill-conditioned inputs (cond 1e12)
    and the Householder fallback on exact rank deficiency."""
    rng = np.random.default_rng(4)
    mats = [rng.standard_normal((300, 40)), _decaying(rng, 500, 69, 0.67)[0], rng.standard_normal((400, 150))]
    deficient = rng.standard_normal((200, 30))
    deficient[:, 5:10] = 0.0
    mats.append(deficient)
    Q, R = bl.cholqr(_pack(mats), want_r=True)
    for a, q, r in zip(mats, _unpack(Q), _unpack(R), strict=True):
        assert np.abs(q.T @ q - np.eye(q.shape[1])).max() < 1e-12
        assert np.abs(np.tril(r, -1)).max() < 1e-12 * np.abs(r).max()
        assert np.abs(q @ r - a).max() < 1e-12 * np.abs(a).max()


def test_cholqr_one_call_matches_launch_by_launch(backend, monkeypatch):
    """The one-call CholeskyQR (fs_cholqr_f64) issues the launches of the
    launch-by-launch path (split-K Gram, factor-only Cholesky, blocked inverse,
    Y R^-1, running product of the factors) with the same descriptors: on the
    device Q, R and every pass's status words are bit-identical, for 1, 2 and 3
    passes, with and without R, on packed and transposed (strided) inputs whose
    Gram products split into 1 to 16 parts, including an all-zero block that
    breaks the factorisation down (its Q is NaN on both paths, its R unspecified)."""
    import torch

    rng = np.random.default_rng(11)
    shapes = ((5, 1), (64, 64), (250, 37), (3100, 20), (4000, 128), (130, 90))
    mats = [rng.standard_normal(sh) for sh in shapes]
    deficient = rng.standard_normal((200, 30))
    deficient[:, 5:10] = 0.0
    mats += [deficient, np.zeros((120, 16))]
    inputs = [_pack(mats), _pack([rng.standard_normal((40, 300)), rng.standard_normal((7, 200))]).T()]
    exact = backend == "gpu"
    for Y in inputs:
        for passes, want_r in ((1, False), (2, False), (2, True), (3, False), (3, True)):
            res = []
            for native in (True, False):
                monkeypatch.setattr(bl, "NATIVE_CHOLQR", native)
                statuses = []
                out = bl.cholqr(Y, passes=passes, want_r=want_r, statuses=statuses)
                Q, R = out if want_r else (out, None)
                res.append((_unpack(Q), None if R is None else _unpack(R), _ops.download(torch.stack(statuses))))
            (q1, r1, s1), (q2, r2, s2) = res
            tag = (len(Y), passes, want_r)
            assert np.array_equal(s1, s2), tag
            # a breakdown leaves Q all NaN and R unspecified
            broken = s1.reshape(passes, len(Y)).any(axis=0)
            for i in np.nonzero(broken)[0]:
                assert np.isnan(q1[i]).all() and np.isnan(q2[i]).all(), tag
            pairs = [(a, b) for i, (a, b) in enumerate(zip(q1, q2, strict=True)) if not broken[i]]
            if want_r:
                pairs += [(a, b) for i, (a, b) in enumerate(zip(r1, r2, strict=True)) if not broken[i]]
            for a, b in pairs:
                if exact:
                    assert np.array_equal(a, b, equal_nan=True), tag
                else:
                    ok = np.isfinite(a)
                    assert np.array_equal(ok, np.isfinite(b)), tag
                    assert np.abs(a[ok] - b[ok]).max(initial=0.0) < 1e-12 * max(1.0, np.abs(a[ok]).max(initial=0.0)), tag
    # read-back mode: the zero block is redone with Householder QR on both paths
    for native in (True, False):
        monkeypatch.setattr(bl, "NATIVE_CHOLQR", native)
        Q, R = bl.cholqr(inputs[0], want_r=True)
        for a, q, r in zip(mats, _unpack(Q), _unpack(R), strict=True):
            assert np.abs(q.T @ q - np.eye(q.shape[1])).max() < 1e-12
            assert np.abs(q @ r - a).max() <= 1e-12 * max(1.0, np.abs(a).max())


def test_rsvd_recovers_from_cholesky_breakdown(backend, monkeypatch):
    """An all-zero block breaks CholeskyQR down (zero pivot even with the shift);
    the deferred status check redoes it with Householder QR and splices the
    result in. The rank-deficient and healthy blocks next to it keep their
    CholeskyQR results."""
    rng = np.random.default_rng(6)
    low = rng.standard_normal((200, 5)) @ rng.standard_normal((5, 150))
    good, spec = _decaying(rng, 180, 160, 0.7)
    zero = np.zeros((120, 90))
    spliced = []
    orig = bl._splice_rsvd
    monkeypatch.setattr(bl, "_splice_rsvd", lambda A, ks, failed, *a: (spliced.append(list(failed)), orig(A, ks, failed, *a)))
    U, S, s_off, s_list, VT = bl.rsvd(_pack([low, good, zero]), [20, 20, 20], seed=2)
    assert spliced == [[2]]
    ref_low = np.linalg.svd(low, compute_uv=False)
    assert np.abs(s_list[0][:5] - ref_low[:5]).max() < 1e-12 * ref_low[0]
    assert np.abs(s_list[1][:20] - spec[:20]).max() < 1e-10
    assert np.abs(s_list[2]).max() == 0.0
    us, vts = _unpack(U), _unpack(VT)
    assert np.isfinite(us[2]).all() and np.isfinite(vts[2]).all()
    for i, a in enumerate([low, good]):
        rec = (us[i][:, :5] * s_list[i][:5]) @ vts[i][:5]
        ref = np.linalg.svd(a)
        assert np.abs(rec - (ref[0][:, :5] * ref[1][:5]) @ ref[2][:5]).max() < 1e-10 * ref[1][0]


def test_triangular_inverse_blocked(backend):
    """fs_batched_triinv: R^-1 of upper triangular matrices of every order up
    to 128 across the 16-row block boundaries (identity padding), against
    scipy's triangular solve; a nonzero factorisation status gives NaN and a
    factor-only Cholesky followed by the inverse matches the fused form."""
    import torch

    rng = np.random.default_rng(9)
    orders = [1, 3, 15, 16, 17, 31, 33, 69, 90, 127, 128]
    Rs = []
    for l in orders:
        r = np.triu(rng.standard_normal((l, l))) / np.sqrt(l)
        r[np.diag_indices(l)] = 1.0 + rng.random(l)
        Rs.append(r)
    sizes = np.array([l * l for l in orders])
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    R = _ops.upload(np.concatenate([r.reshape(-1) for r in Rs]))
    X = _ops.upload(np.full(int(sizes.sum()), 7.0))
    status = _ops.upload(np.array([0] * (len(orders) - 1) + [3]), dtype=torch.int32)
    _ops.batched_triinv({"l": np.array(orders), "r_off": offs, "x_off": offs}, R, X, status)
    x = _ops.download(X)
    for i, (l, r) in enumerate(zip(orders, Rs, strict=True)):
        got = x[offs[i] : offs[i] + l * l].reshape(l, l)
        if i == len(orders) - 1:
            assert np.isnan(got).all()
            continue
        ref = lg.solve_triangular(r, np.eye(l), lower=False)
        assert np.abs(got - ref).max() < 1e-13 * max(1.0, np.abs(ref).max()), l
        assert np.abs(np.tril(got, -1)).max() == 0.0
    # factor only + blocked inverse against the fused Cholesky + inverse
    l, n = 90, 3
    Gs = []
    for _ in range(n):
        y = rng.standard_normal((300, l))
        Gs.append((y.T @ y).reshape(-1))
    G = _ops.upload(np.concatenate(Gs))
    base = {"l": np.full(n, l), "g_off": np.arange(n) * l * l, "g_stride": l * l, "n_part": 1,
            "r_off": np.arange(n) * l * l, "rinv_off": np.arange(n) * l * l, "shift_rel": 0.0}
    R1, X1, R2, X2 = (_ops.empty(n * l * l) for _ in range(4))
    st1, st2 = (_ops.zeros(n, dtype=torch.int32) for _ in range(2))
    _ops.batched_cholinv(base, G, R1, X1, st1)
    _ops.batched_cholinv(base, G, R2, None, st2)
    _ops.batched_triinv({"l": np.full(n, l), "r_off": base["r_off"], "x_off": base["r_off"]}, R2, X2, st2)
    assert (_ops.download(st1) == 0).all() and (_ops.download(st2) == 0).all()
    r1, r2 = _ops.download(R1), _ops.download(R2)
    assert np.abs(r1 - r2).max() < 1e-13 * np.abs(r1).max()
    x1, x2 = _ops.download(X1), _ops.download(X2)
    assert np.abs(x1 - x2).max() < 1e-12 * np.abs(x1).max()


@pytest.mark.parametrize(
    "orders",
    [[16, 1, 3], [31, 5, 17], [40, 2, 33], [64, 9, 49], [80, 65, 17], [96, 1, 90], [112, 50, 97], [128, 7, 100]],
)
def test_cholesky_and_inverse_register_tiles(backend, orders):
    """The blocked Cholesky and triangular inverse size their register tiles to
    the batch's largest order: one batch per tile size, mixed orders in it, the
    factor (factor-only and fused forms) and the inverse against scipy."""
    import torch

    rng = np.random.default_rng(sum(orders))
    n = len(orders)
    sq = np.array([l * l for l in orders])
    offs = np.concatenate([[0], np.cumsum(sq)[:-1]])
    Gs, Rref = [], []
    for l in orders:
        y = rng.standard_normal((l + 30, l))
        g = y.T @ y
        Gs.append(g.reshape(-1))
        Rref.append(lg.cholesky(g, lower=False))
    G = _ops.upload(np.concatenate(Gs))
    desc = {"l": np.array(orders), "g_off": offs, "g_stride": sq, "n_part": 1, "r_off": offs,
            "rinv_off": offs, "shift_rel": 0.0}
    R1, R2, X1, X2 = (_ops.empty(int(sq.sum())) for _ in range(4))
    st1, st2 = (_ops.zeros(n, dtype=torch.int32) for _ in range(2))
    _ops.batched_cholinv(desc, G, R1, None, st1)  # factor only
    _ops.batched_cholinv(desc, G, R2, X1, st2)  # fused factor + inverse
    _ops.batched_triinv({"l": np.array(orders), "r_off": offs, "x_off": offs}, R1, X2, st1)
    assert (_ops.download(st1) == 0).all() and (_ops.download(st2) == 0).all()
    r1, r2, x1, x2 = (_ops.download(t) for t in (R1, R2, X1, X2))
    for i, l in enumerate(orders):
        sl = slice(offs[i], offs[i] + l * l)
        ref = Rref[i]
        inv = lg.solve_triangular(ref, np.eye(l), lower=False)
        for got in (r1[sl], r2[sl]):
            assert np.abs(got.reshape(l, l) - ref).max() < 1e-12 * np.abs(ref).max(), (orders, l)
        for got in (x1[sl], x2[sl]):
            assert np.abs(got.reshape(l, l) - inv).max() < 1e-11 * np.abs(inv).max(), (orders, l)


def test_factor_only_cholesky(backend):
    """Factor-only batched Cholesky (the four-columns-per-barrier kernel without
    its inverse phase) against scipy: orders across block boundaries, Gram matrices
    given as split-K partial sums, the relative shift, and the breakdown status
    of a matrix that is not positive definite."""
    import torch

    rng = np.random.default_rng(12)
    orders = [1, 5, 15, 16, 17, 33, 69, 90, 127, 128]
    nparts, shift_rel = 3, 1e-3
    G_parts, refs, descs = [], [], {"l": [], "g_off": [], "g_stride": [], "n_part": [], "r_off": [],
                                    "rinv_off": [], "shift_rel": []}
    g_off = r_off = 0
    for l in orders:
        ys = [rng.standard_normal((40 + l, l)) for _ in range(nparts)]
        parts = [y.T @ y for y in ys]
        G_parts.extend(p.reshape(-1) for p in parts)
        g = sum(parts)
        refs.append(lg.cholesky(g + shift_rel * np.trace(g) * np.eye(l), lower=False))
        for k, v in (("l", l), ("g_off", g_off), ("g_stride", l * l), ("n_part", nparts), ("r_off", r_off),
                     ("rinv_off", 0), ("shift_rel", shift_rel)):
            descs[k].append(v)
        g_off += nparts * l * l
        r_off += l * l
    # a symmetric matrix with a negative eigenvalue breaks down
    bad = np.diag(np.concatenate([np.ones(20), [-1.0], np.ones(19)]))
    G_parts.append(bad.reshape(-1))
    for k, v in (("l", 40), ("g_off", g_off), ("g_stride", 1600), ("n_part", 1), ("r_off", r_off),
                 ("rinv_off", 0), ("shift_rel", 0.0)):
        descs[k].append(v)
    G = _ops.upload(np.concatenate(G_parts))
    R = _ops.upload(np.zeros(r_off + 1600))
    status = _ops.upload(np.full(len(orders) + 1, -7), dtype=torch.int32)
    _ops.batched_cholinv({k: np.array(v) for k, v in descs.items()}, G, R, None, status)
    st = _ops.download(status)
    # the device reports the failing pivot (index 20, plus one); the emulation only that it failed
    assert (st[:-1] == 0).all() and st[-1] > 0 and (backend != "gpu" or st[-1] == 21)
    r = _ops.download(R)
    for l, ref, off in zip(orders, refs, descs["r_off"], strict=False):
        got = r[off : off + l * l].reshape(l, l)
        assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max(), l
        assert np.abs(np.tril(got, -1)).max() == 0.0
