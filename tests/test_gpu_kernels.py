"""Native kernels on the GPU against numpy/dense oracles (bit-exact for data
movement, ~1e-13 relative for FP64 arithmetic)."""

import numpy as np
import pytest
import torch

from frostspin_amd import _ops

pytestmark = pytest.mark.gpu


def _dev(a):
    return _ops.upload(np.asarray(a, dtype=np.float64))


GEMM_SIZES_SMALL = [(1, 1, 1), (3, 70, 5), (64, 64, 16), (65, 63, 17), (130, 7, 200), (17, 129, 33)]
# one group with >= 256 64x64 tiles switches the launch to the 64x64 kernel
GEMM_SIZES_LARGE = [(1100, 1050, 70), (33, 47, 65), (70, 1, 130)]
# thin launches (every group at most 96 wide on one side, long K: the
# range-finder and projector products), ragged groups, split over K
GEMM_SIZES_THIN_N = [(1444, 46, 1444), (700, 40, 900), (150, 20, 300), (9, 3, 260)]
GEMM_SIZES_THIN_N96 = [(1500, 91, 600), (600, 70, 700), (130, 91, 256)]
GEMM_SIZES_THIN_M = [(46, 1444, 1444), (69, 800, 700), (12, 300, 500)]
GEMM_SIZES_THIN_M80 = [(80, 900, 400), (77, 300, 300)]


@pytest.mark.parametrize(
    "sizes",
    [GEMM_SIZES_SMALL, GEMM_SIZES_LARGE, GEMM_SIZES_THIN_N, GEMM_SIZES_THIN_N96, GEMM_SIZES_THIN_M, GEMM_SIZES_THIN_M80],
    ids=["tile32", "tile64", "thin_n48", "thin_n96", "thin_m80", "thin_m80b"],
)
def test_mfma_gemm_layouts_and_uneven_groups(gpu, sizes):
    rng = np.random.default_rng(0)
    for a_f in (False, True):
        for b_f in (False, True):
            A_parts, B_parts, groups, refs = [], [], [], []
            ao = bo = co = 0
            for m, n, k in sizes:
                a = rng.standard_normal((m, k))
                b = rng.standard_normal((k, n))
                A_parts.append((a.T if a_f else a).reshape(-1))
                B_parts.append((b.T if b_f else b).reshape(-1))
                ars, acs = (1, m) if a_f else (k, 1)
                brs, bcs = (1, k) if b_f else (n, 1)
                groups.append((m, n, k, ao, ars, acs, bo, brs, bcs, co, n))
                refs.append((co, a @ b))
                ao += m * k
                bo += k * n
                co += m * n
            A = _dev(np.concatenate(A_parts))
            B = _dev(np.concatenate(B_parts))
            C = _ops.upload(np.full(co, 7.0))
            _ops.grouped_gemm(groups, A, B, C, alpha=2.0, beta=0.5)
            c = _ops.download(C)
            for (off, ref), (m, n, _) in zip(refs, sizes, strict=True):
                got = c[off : off + m * n].reshape(m, n)
                expect = 2.0 * ref + 3.5
                err = float(np.abs(got - expect).max() / max(1.0, np.abs(expect).max()))
                assert err <= 1e-12, f"group {(m, n)} a_f={a_f} b_f={b_f} rel err {err:.2e}"


# (m, n, k) of the wide-product kernel family: few rows and a short K over a
# long n (a site tensor applied to an environment tensor), each row-count
# instance (16, 32, 64, 128 rows), K chunks of 32 with a ragged last chunk,
# column counts that end inside a workgroup
GEMM_SIZES_WIDE = [(5, 3000, 7), (16, 300, 128), (30, 5000, 12), (36, 20001, 12), (64, 9000, 34),
                   (116, 20000, 34), (34, 20000, 116), (128, 8192, 128), (100, 2100, 40), (1, 4096, 1)]


def test_wide_gemm_instances(gpu):
    rng = np.random.default_rng(5)
    for a_f in (False, True):
        for b_f in (False, True):  # column-major B goes to the tile kernel
            A_parts, B_parts, groups, refs = [], [], [], []
            ao = bo = co = 0
            for m, n, k in GEMM_SIZES_WIDE:
                a = rng.standard_normal((m, k))
                b = rng.standard_normal((k, n))
                A_parts.append((a.T if a_f else a).reshape(-1))
                B_parts.append((b.T if b_f else b).reshape(-1))
                ars, acs = (1, m) if a_f else (k, 1)
                brs, bcs = (1, k) if b_f else (n, 1)
                groups.append((m, n, k, ao, ars, acs, bo, brs, bcs, co, n))
                refs.append((co, a @ b))
                ao += m * k
                bo += k * n
                co += m * n
            A = _dev(np.concatenate(A_parts))
            B = _dev(np.concatenate(B_parts))
            C = _ops.upload(np.full(co, -3.0))
            _ops.grouped_gemm(groups, A, B, C, alpha=-1.5, beta=2.0)
            c = _ops.download(C)
            for (off, ref), (m, n, _) in zip(refs, GEMM_SIZES_WIDE, strict=True):
                got = c[off : off + m * n].reshape(m, n)
                expect = -1.5 * ref - 6.0
                err = float(np.abs(got - expect).max() / max(1.0, np.abs(expect).max()))
                assert err <= 1e-12, f"group {(m, n)} a_f={a_f} b_f={b_f} rel err {err:.2e}"


def test_gemm_many_groups_chunked(gpu):
    rng = np.random.default_rng(1)
    groups, refs = [], []
    A_parts, B_parts = [], []
    ao = bo = co = 0
    for _ in range(75):  # > 32 groups per launch
        m, n, k = rng.integers(1, 40, size=3)
        a = rng.standard_normal((m, k))
        b = rng.standard_normal((k, n))
        A_parts.append(a.reshape(-1))
        B_parts.append(b.reshape(-1))
        groups.append((m, n, k, ao, k, 1, bo, n, 1, co, n))
        refs.append((co, m, n, a @ b))
        ao += m * k
        bo += k * n
        co += m * n
    C = _ops.empty(co)
    _ops.grouped_gemm(groups, _dev(np.concatenate(A_parts)), _dev(np.concatenate(B_parts)), C)
    c = _ops.download(C)
    for off, m, n, ref in refs:
        err = float(np.abs(c[off : off + m * n].reshape(m, n) - ref).max())
        assert err < 1e-12, f"group {(m, n)} err {err:.2e}"


def test_split_gemm_precision(gpu):
    """Compensated FP32 form: x = hi + lo (both FP32, x - hi - lo <= u^2 |x|),
    the three products hi*hi + hi*lo + lo*hi (lo*lo dropped, <= u^2 |a||b|) on
    FP32 MFMA with FP32 accumulation over one 16-deep K chunk (12 roundings of
    the running chunk sum), FP64 across chunks. Guaranteed componentwise bound:
    |C - AB| <= 16 u (|A| |B|), u = 2^-24; checked on every entry against FP64
    products summed by math.fsum (their own error, ~1e-16 relative, is far
    below the bound)."""
    import math

    rng = np.random.default_rng(2)
    m, n, k = 96, 80, 300
    a = rng.standard_normal((m, k))
    b = rng.standard_normal((k, n))
    C = _ops.empty(m * n)
    _ops.grouped_gemm([(m, n, k, 0, k, 1, 0, n, 1, 0, n)], _dev(a), _dev(b), C, split=True)
    got = _ops.download(C).reshape(m, n)
    exact = np.array([[math.fsum(a[i] * b[:, j]) for j in range(n)] for i in range(m)])  # FP64 products
    bound = 16 * 2.0**-24 * (np.abs(a) @ np.abs(b))
    assert (np.abs(got - exact) <= bound).all(), float((np.abs(got - exact) / bound).max())
    # and it is not better than FP32 accumulation allows: the FP64 path is ~1e-15
    err = float(np.abs(got - exact).max() / np.abs(exact).max())
    assert err > 1e-10, f"split gemm rel err {err:.2e} suspiciously small"


def _graded_lower(rng, n, decay):
    """Transposed R factor of a matrix with singular values exp(-decay k / n):
    the lower-triangular small matrix the randomized SVD hands to the Jacobi SVD."""
    s = np.exp(-np.arange(n) * (decay / n))
    u, _ = np.linalg.qr(rng.standard_normal((n, n)))
    v, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return np.ascontiguousarray(np.linalg.qr(((u * s) @ v.T).T, mode="r").T)


def _check_batched_svd(mats):
    parts, descs = [], []
    ao = uo = so = vo = wo = 0
    for a in mats:
        m, n = a.shape
        q = min(m, n)
        parts.append(a.reshape(-1))
        descs.append((m, n, ao, n, 1, uo, so, vo, wo))
        ao += m * n
        uo += m * q
        so += q
        vo += q * n
        wo += q * (max(m, n) + q)
    A = _dev(np.concatenate(parts))
    U, S, VT, W = _ops.empty(uo), _ops.empty(so), _ops.empty(vo), _ops.empty(wo)
    _ops.batched_svd(descs, A, U, S, VT, W)
    u, s, vt = _ops.download(U), _ops.download(S), _ops.download(VT)
    for (m, n, _, _, _, uo_, so_, vo_, _), a in zip(descs, mats, strict=True):
        q = min(m, n)
        uu = u[uo_ : uo_ + m * q].reshape(m, q)
        ss = s[so_ : so_ + q]
        vv = vt[vo_ : vo_ + q * n].reshape(q, n)
        ref = np.linalg.svd(a, compute_uv=False)
        tag = f"shape {(m, n)}"
        assert np.all(np.diff(ss) <= 0), tag
        e_s = float(np.abs(ss - ref).max() / ref[0])
        e_r = float(np.abs(uu * ss @ vv - a).max() / ref[0])
        nz = ss > 1e-10 * ss[0]
        e_u = float(np.abs(uu[:, nz].T @ uu[:, nz] - np.eye(nz.sum())).max())
        e_v = float(np.abs(vv @ vv.T - np.eye(q)).max())
        assert e_s < 1e-13 and e_r < 1e-12 and e_u < 1e-12 and e_v < 1e-12, (tag, e_s, e_r, e_u, e_v)


def test_jacobi_svd_batched(gpu):
    # p <= 128 in LDS takes the register form, (20, 130) the strided form in
    # LDS and (200, 200) the strided form in the work buffer, all in one launch
    rng = np.random.default_rng(3)
    shapes = [(1, 1), (5, 3), (3, 5), (40, 40), (100, 37), (20, 130), (200, 200)]
    mats = []
    for m, n in shapes:
        a = rng.standard_normal((m, n))
        if m == 40:
            a[:, 5] = a[:, 6]  # rank deficient
        mats.append(a)
    _check_batched_svd(mats)


@pytest.mark.parametrize(
    "shapes",
    [
        [(24, 24), (7, 31), (2, 2)],  # register length 2 (p <= 32)
        [(45, 45), (64, 50), (33, 17)],  # 4 (p <= 64)
        [(70, 70), (96, 12), (65, 65)],  # 6 (p <= 96)
        [(128, 128), (127, 3), (97, 97)],  # 8 (p <= 128)
    ],
)
def test_jacobi_svd_register_lengths(gpu, shapes):
    rng = np.random.default_rng(len(shapes) + shapes[0][0])
    mats = []
    for m, n in shapes:
        if m == n:
            a = _graded_lower(rng, m, 20.0)
            if m > 40:
                a[:, 3] = a[:, 4]  # rank deficient
        else:
            a = rng.standard_normal((m, n))
        mats.append(a)
    _check_batched_svd(mats)


def test_norm_trace_copy_fill(gpu):
    rng = np.random.default_rng(4)
    x = rng.standard_normal(100_003)
    X = _dev(x)
    seg = np.array([0, 10, 5000, 100_003])
    w = np.array([1.0, 3.0, 2.0])
    got = float(_ops.download(_ops.weighted_norm2(X, seg, w))[0])
    ref = sum(wi * np.sum(x[seg[i] : seg[i + 1]] ** 2) for i, wi in enumerate(w))
    assert abs(got - ref) <= 1e-13 * ref
    m = rng.standard_normal((7, 7))
    tr = float(_ops.download(_ops.weighted_trace(_dev(m), [0], [7], [7], [2.0]))[0])
    assert abs(tr - 2 * np.trace(m)) < 1e-13
    src = rng.standard_normal((6, 9))
    dst = _ops.zeros(54)
    cw = _dev(np.arange(9.0))
    idx = _ops.upload(np.array([8, 0, 3], dtype=np.int64), dtype=torch.int64)
    _ops.copy2d([(6, 3, 0, 9, 1, 0, 3, 1, 0, 2, 0, 1.5, 0.0)], _dev(src), dst, col_w=cw, col_idx=idx)
    got = _ops.download(dst)[:18].reshape(6, 3)
    assert np.array_equal(got, 1.5 * src[:, [8, 0, 3]] * np.arange(2.0, 5.0))
    y = _ops.empty(1001)
    _ops.fill_normal_(y, 5)
    yy = _ops.download(y)
    assert abs(yy.mean()) < 0.15 and abs(yy.std() - 1) < 0.1
    y2 = _ops.empty(1001)
    _ops.fill_normal_(y2, 5)
    assert np.array_equal(yy, _ops.download(y2))


@pytest.mark.parametrize("sizes", [GEMM_SIZES_SMALL, GEMM_SIZES_LARGE], ids=["tile32", "tile64"])
def test_mfma_gemm_per_group_operand_bases(gpu, sizes):
    """fs_grouped_gemm_f64_multi: every group reads and writes its own buffers
    (blocks of different tensors in one launch); 40 groups cover two dispatches."""
    rng = np.random.default_rng(3)
    shapes = (list(sizes) * 14)[:40]
    As, Bs, Cs, groups, refs = [], [], [], [], []
    for i, (m, n, k) in enumerate(shapes):
        a = rng.standard_normal((m, k))
        b = rng.standard_normal((k, n))
        pad = int(rng.integers(0, 5))  # operands at nonzero offsets inside their buffers
        As.append(_dev(np.concatenate((np.zeros(pad), a.T.reshape(-1) if i % 2 else a.reshape(-1)))))
        Bs.append(_dev(np.concatenate((np.zeros(pad), b.reshape(-1)))))
        Cs.append(_ops.upload(np.full(pad + m * n, 7.0)))
        ars, acs = (1, m) if i % 2 else (k, 1)
        groups.append((m, n, k, pad, ars, acs, pad, n, 1, pad, n))
        refs.append((pad, a @ b))
    _ops.grouped_gemm(groups, As, Bs, Cs, alpha=2.0, beta=0.5)
    for (pad, ref), C, (m, n, _) in zip(refs, Cs, shapes, strict=True):
        c = _ops.download(C)
        assert np.all(c[:pad] == 7.0)
        expect = 2.0 * ref + 3.5
        err = float(np.abs(c[pad:].reshape(m, n) - expect).max() / max(1.0, np.abs(expect).max()))
        assert err <= 1e-12


def _check_groups(shapes, layouts, rng, alpha=1.5, beta=0.25):
    """Grouped GEMM of the given (m, n, k) shapes and (A col-major, B col-major)
    layouts against numpy, with alpha / beta."""
    A_parts, B_parts, groups, refs = [], [], [], []
    ao = bo = co = 0
    for (m, n, k), (a_f, b_f) in zip(shapes, layouts, strict=True):
        a = rng.standard_normal((m, k))
        b = rng.standard_normal((k, n))
        A_parts.append((a.T if a_f else a).reshape(-1))
        B_parts.append((b.T if b_f else b).reshape(-1))
        ars, acs = (1, m) if a_f else (k, 1)
        brs, bcs = (1, k) if b_f else (n, 1)
        groups.append((m, n, k, ao, ars, acs, bo, brs, bcs, co, n))
        refs.append((co, a @ b))
        ao += m * k
        bo += k * n
        co += m * n
    C0 = rng.standard_normal(co)
    C = _ops.upload(C0)
    _ops.grouped_gemm(groups, _dev(np.concatenate(A_parts)), _dev(np.concatenate(B_parts)), C, alpha=alpha, beta=beta)
    c = _ops.download(C)
    for (off, ref), (m, n, _) in zip(refs, shapes, strict=True):
        expect = alpha * ref + beta * C0[off : off + m * n].reshape(m, n)
        got = c[off : off + m * n].reshape(m, n)
        err = float(np.abs(got - expect).max() / max(1.0, np.abs(expect).max()))
        assert err <= 1e-12, f"shape {(m, n)} rel err {err:.2e}"


def test_gemm_skinny_rows(gpu):
    """m <= 16, k <= 128, n >= 256: the streaming kernel (both B layouts, mixed
    with tile-kernel groups in one call)."""
    rng = np.random.default_rng(11)
    shapes = [(12, 5000, 36), (1, 300, 1), (16, 777, 128), (3, 2049, 7), (40, 300, 20), (12, 100, 36)]
    layouts = [(False, False), (True, True), (False, True), (True, False), (False, False), (True, False)]
    _check_groups(shapes, layouts, rng)


def test_gemm_split_k(gpu):
    """Few output tiles over a long K: split over K, partials reduced in order;
    repeated calls give bit-identical results."""
    rng = np.random.default_rng(12)
    shapes = [(40, 50, 6000), (64, 30, 2500), (17, 70, 1025), (5, 5, 700)]
    layouts = [(False, False), (True, False), (False, True), (True, True)]
    _check_groups(shapes, layouts, rng)
    a = _dev(rng.standard_normal(60 * 4000))
    b = _dev(rng.standard_normal(4000 * 50))
    outs = []
    for _ in range(2):
        C = _ops.empty(60 * 50)
        _ops.grouped_gemm([(60, 50, 4000, 0, 4000, 1, 0, 50, 1, 0, 50)], a, b, C)
        outs.append(_ops.download(C))
    assert np.array_equal(outs[0], outs[1])


def test_gemm_split_k_fills_slots(gpu):
    """Range-finder products B = Q^T M (l x n over K = m, l << n): 167 tiles of
    32 x 32 over K up to 1444 are split to fill the resident workgroup slots
    (one tile kernel and one reduction), and match numpy for every layout."""
    rng = np.random.default_rng(13)
    shapes = [(19, 660, 668), (36, 1432, 1444), (21, 1216, 1220), (4, 576, 576)]
    layouts = [(True, False), (True, False), (False, True), (False, False)]
    _check_groups(shapes, layouts, rng)
    assert _ops._gemm_dispatches() == 2


def test_gemm_split_k_shares_per_group(gpu):
    """Range-finder launch of ten sectors (m = k from 164 to 1412, 20 to 66
    columns): the K panels are dealt per group, so the long-K sectors run more
    splits than the short ones in the same launch (one tile kernel and one
    reduction); against numpy for two layouts, and bit-identical when the same
    call is repeated."""
    rng = np.random.default_rng(16)
    ms = [164, 300, 420, 560, 700, 840, 980, 1120, 1260, 1412]
    ns = [20, 26, 30, 36, 40, 46, 50, 56, 60, 66]
    shapes = [(m, n, m) for m, n in zip(ms, ns, strict=True)]
    _check_groups(shapes, [(False, False)] * 10, rng)
    assert _ops._gemm_dispatches() == 2
    _check_groups(shapes, [(i % 2 == 1, False) for i in range(10)], rng)
    groups, ao, bo, co = [], 0, 0, 0
    for m, n, k in shapes:
        groups.append((m, n, k, ao, k, 1, bo, n, 1, co, n))
        ao, bo, co = ao + m * k, bo + k * n, co + m * n
    a = _dev(rng.standard_normal(ao))
    b = _dev(rng.standard_normal(bo))
    outs = []
    for _ in range(2):
        C = _ops.empty(co)
        _ops.grouped_gemm(groups, a, b, C)
        outs.append(_ops.download(C))
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize(
    "shapes",
    [
        # thin n: the range-finder products A X, uneven groups
        [(1444, 69, 1444), (668, 43, 668), (200, 20, 150), (37, 5, 3)],
        [(600, 128, 300), (129, 100, 50)],
        [(300, 20, 100), (129, 32, 40), (7, 1, 1)],
        # thin m: the projections X^T A
        [(69, 1444, 1444), (43, 668, 668), (20, 200, 150), (5, 37, 3)],
        [(128, 700, 200), (100, 600, 64)],
    ],
    ids=["n69", "n128", "n32", "m69", "m128"],
)
def test_gemm_thin_shapes(gpu, shapes):
    """Products with one thin side over long K ranges (32 x 32 tiles, split-K,
    four K panels in flight, K tails of every length modulo the panel depth):
    every operand layout, alpha and beta, uneven groups, against numpy."""
    rng = np.random.default_rng(14)
    for layouts in ([(False, False)] * len(shapes), [(True, True)] * len(shapes),
                    [(i % 2 == 0, i % 2 == 1) for i in range(len(shapes))],
                    [(i % 2 == 1, i % 2 == 0) for i in range(len(shapes))]):
        _check_groups(shapes, layouts, rng)


@pytest.mark.parametrize("k", [0, 1, 15, 16, 17, 48, 63, 64, 65, 80, 129, 1000])
def test_gemm_k_tails(gpu, k):
    """Every K length modulo the panel (16) and the panel depth (4 x 16), K = 0
    (C = beta C) included, with and without split-K: one small and one large
    launch per K against numpy."""
    rng = np.random.default_rng(15 + k)
    # a (2, 2, 2) group keeps the operand buffers non-empty when k = 0
    _check_groups([(40, 30, k), (3, 50, k), (2, 2, 2)], [(False, True), (True, False), (False, False)], rng)
    _check_groups([(300, 260, k), (2, 2, 2)], [(False, False), (True, True)], rng)
