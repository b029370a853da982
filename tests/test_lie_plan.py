"""Structural-transform plans (fs_lie_plan_*) on synthetic task tables: every
kernel path (register form n_in <= 32, wide four-wave form 32 < n_in <= 128,
streamed form above) against the numpy restatement of the plan semantics, and
the restatement itself against a dense matrix built element by element.

Each task reads n_in rows and writes n_out rows of E = prod(radices) external
elements; the rows of one task interleave in their buffer (row stride n * C),
absent input rows (base -1) contribute zero, and the digit order of the output
is an independent permutation of the input's, as for a permute."""

import math

import numpy as np
import pytest
import torch

from oracle.device_emulator import EmulatedLiePlan

# (n_in, n_out, radices): E not a multiple of 64, E < 64, one n_in per path
SHAPES = [
    (1, 1, (2, 3, 5)),
    (3, 5, (7, 2)),
    (6, 6, (2, 2, 2, 2, 2, 2, 15, 3)),
    (21, 21, (4, 3, 5)),
    (32, 32, (3, 3, 3)),
    (33, 40, (2, 5, 7)),
    (51, 51, (2, 2, 2, 2, 2, 2, 3)),
    (70, 70, (5, 13)),
    (128, 9, (11,)),
    (40, 130, (3, 2, 5)),
    (5, 70, (3, 7)),
    (36, 2, (64,)),
    (130, 17, (2, 31)),
    (0, 3, (4,)),
]


def _layout(rng, radices):
    """Random split of the digits into row and column digits of a block, in a
    random order: per-digit strides (row, column) and the block shape."""
    nd = len(radices)
    perm = rng.permutation(nd)
    split = int(rng.integers(0, nd + 1))
    r_stride = np.zeros(nd, dtype=np.int64)
    c_stride = np.zeros(nd, dtype=np.int64)
    for part, stride in ((perm[:split], r_stride), (perm[split:], c_stride)):
        s = 1
        for k in part[::-1]:
            stride[k] = s
            s *= radices[k]
    R = math.prod(radices[k] for k in perm[:split])
    C = math.prod(radices[k] for k in perm[split:])
    return R, C, r_stride, c_stride


def synthetic_plan(rng, shapes, absent=0.15):
    tasks, rows, digits, w = [], [], [], []
    src_off = dst_off = 0
    for n_in, n_out, radices in shapes:
        E = math.prod(radices)
        Ri, Ci, ir, ic = _layout(rng, radices)
        Ro, Co, orr, oc = _layout(rng, radices)
        row0 = len(rows)
        for i in range(n_in):
            if rng.random() < absent:
                rows.append((-1, 0, 0))
            else:
                rows.append((src_off + i * Ci, n_in * Ci, 1))
        src_off += n_in * Ri * Ci
        for o in range(n_out):
            rows.append((dst_off + o * Co, n_out * Co, 1))
        dst_off += n_out * Ro * Co
        pad = (n_out + 7) // 8 * 8
        w_off = len(w)
        w.extend(rng.standard_normal(n_in * pad))
        dig0 = len(digits)
        for k, radix in enumerate(radices):
            digits.append((radix, ir[k], ic[k], orr[k], oc[k]))
        tasks.append((E, w_off, n_in, n_out, pad, row0, len(radices), dig0))
    return (np.array(tasks, dtype=np.int64), np.array(rows, dtype=np.int64).reshape(-1, 3),
            np.array(digits, dtype=np.int64).reshape(-1, 5), np.array(w), src_off, dst_off)


def _dense_reference(tasks, rows, digits, w, src, dst_size):
    """Element by element: output row o of task t at external index e is
    sum_i W^T[i, o] * input row i at e (mixed radix, last digit fastest)."""
    out = np.zeros(dst_size)
    for E, w_off, n_in, n_out, pad, row0, ndig, dig0 in tasks:
        Wt = w[w_off : w_off + n_in * pad].reshape(n_in, pad) if n_in else np.zeros((0, pad))
        for e in range(E):
            rem, er_i, ec_i, er_o, ec_o = e, 0, 0, 0, 0
            for radix, in_r, in_c, out_r, out_c in digits[dig0 : dig0 + ndig][::-1]:
                d = rem % radix
                rem //= radix
                er_i += d * in_r
                ec_i += d * in_c
                er_o += d * out_r
                ec_o += d * out_c
            x = np.array([src[b + er_i * rs + ec_i * cs] if b >= 0 else 0.0
                          for b, rs, cs in rows[row0 : row0 + n_in]])
            for o in range(n_out):
                b, rs, cs = rows[row0 + n_in + o]
                out[b + er_o * rs + ec_o * cs] = float(Wt[:, o] @ x) if n_in else 0.0
    return out


def test_emulated_plan_matches_elementwise_reference():
    rng = np.random.default_rng(5)
    small = [(1, 1, (2, 3)), (3, 5, (7, 2)), (4, 2, (2, 2, 3)), (0, 2, (3,))]
    tasks, rows, digits, w, ns, nd = synthetic_plan(rng, small)
    src = rng.standard_normal(ns)
    ref = _dense_reference(tasks, rows, digits, w, src, nd)
    dst = torch.full((nd,), np.nan, dtype=torch.float64)
    EmulatedLiePlan(tasks, rows, digits, w, ns, nd).execute(torch.from_numpy(src), dst)
    # every destination element is written exactly once
    assert np.isfinite(dst.numpy()).all()
    np.testing.assert_allclose(dst.numpy(), ref, rtol=1e-13, atol=1e-13)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [0, 1])
def test_device_plan_matches_emulation_on_every_path(gpu, seed):
    from frostspin_amd import _ops

    rng = np.random.default_rng(seed)
    tasks, rows, digits, w, ns, nd = synthetic_plan(rng, SHAPES)
    src = rng.standard_normal(ns)
    ref = torch.full((nd,), np.nan, dtype=torch.float64)
    EmulatedLiePlan(tasks, rows, digits, w, ns, nd).execute(torch.from_numpy(src), ref)
    plan = _ops.LiePlan(tasks, rows, digits, w, ns, nd)
    dst = _ops.upload(np.full(nd, np.nan))
    plan.execute(_ops.upload(src), dst)
    got = _ops.download(dst)
    assert np.isfinite(got).all()
    scale = max(1.0, float(np.abs(ref.numpy()).max()))
    np.testing.assert_allclose(got, ref.numpy(), rtol=0, atol=1e-13 * scale)
    # a second execution of the same plan gives the same bits
    dst2 = _ops.upload(np.full(nd, np.nan))
    plan.execute(_ops.upload(src), dst2)
    assert np.array_equal(_ops.download(dst2), got)


def _numpy_plan_tables(t, axes, nrr_out):
    """The plan tables of t.permute restated with numpy over the cached
    irreps-only structure: the in-block row / column of every input and output
    row from running sums of degeneracy x internal multiplicity per target
    block, digits one per new leg with row-major strides on both sides."""
    from frostspin_amd.symmetric_tensor import lie_group_symmetric_tensor as lg

    reps = t._row_reps + t._col_reps
    ndim, nrr = t._ndim, t._nrr
    irreps_per_leg = tuple(tuple(int(x) for x in r[1]) for r in reps)
    sig = tuple(bool(s) for s in t._signature)
    st = type(t)._permute_structure(irreps_per_leg, sig, nrr, tuple(axes), nrr_out)
    degens = [np.asarray(r[0], dtype=np.int64) for r in reps]
    nsig = st.sig_idx.shape[0]
    D = np.stack([degens[k][st.sig_idx[:, k]] for k in range(ndim)], axis=1)
    edr, edc = lg._outer(degens[:nrr]), lg._outer(degens[nrr:])
    rstart = lg._exclusive_cumsum0(edr[:, None] * st.idr_in)
    cstart = lg._exclusive_cumsum0(edc[:, None] * st.idc_in)
    pos = t._present_map(st.in_targets)
    rs_b, cs_b = t._block_strides()
    sig_e = np.repeat(np.arange(nsig), st.n_in)
    b = pos[st.in_t]
    ir, ic = st.ir_in[sig_e], st.ic_in[sig_e]
    row = rstart[ir, st.in_t] + st.in_a * edr[ir]
    col = cstart[ic, st.in_t] + st.in_b * edc[ic]
    bb = np.maximum(b, 0)
    in_base = np.where(b >= 0, t._boffs[bb] + row * rs_b[bb] + col * cs_b[bb], -1)
    in_rs, in_cs = np.where(b >= 0, rs_b[bb], 0), np.where(b >= 0, cs_b[bb], 0)
    ndeg = [degens[a] for a in axes]
    edr2, edc2 = lg._outer(ndeg[:nrr_out]), lg._outer(ndeg[nrr_out:])
    shapes = np.stack([(edr2[:, None] * st.idr_out).sum(axis=0), (edc2[:, None] * st.idc_out).sum(axis=0)], axis=1)
    noff = lg._exclusive_cumsum0(shapes[:, 0] * shapes[:, 1])
    rstart2 = lg._exclusive_cumsum0(edr2[:, None] * st.idr_out)
    cstart2 = lg._exclusive_cumsum0(edc2[:, None] * st.idc_out)
    sig_o = np.repeat(np.arange(nsig), st.n_out)
    ir2, ic2, t2 = st.ir_out[sig_o], st.ic_out[sig_o], st.out_t
    out_base = (noff[t2] + (rstart2[ir2, t2] + st.out_a * edr2[ir2]) * shapes[t2, 1]
                + cstart2[ic2, t2] + st.out_b * edc2[ic2])
    nrow = st.n_in + st.n_out
    row0 = lg._exclusive_cumsum0(nrow)
    rows = np.zeros((int(nrow.sum()), 3), dtype=np.int64)
    pi = row0[sig_e] + np.arange(sig_e.size) - np.repeat(lg._exclusive_cumsum0(st.n_in), st.n_in)
    rows[pi, 0], rows[pi, 1], rows[pi, 2] = in_base, in_rs, in_cs
    po = row0[sig_o] + st.n_in[sig_o] + np.arange(sig_o.size) - np.repeat(lg._exclusive_cumsum0(st.n_out), st.n_out)
    rows[po, 0], rows[po, 1], rows[po, 2] = out_base, shapes[t2, 1], 1
    Dn = D[:, list(axes)]
    s_old_r, s_old_c = lg._suffix_products(D[:, :nrr]), lg._suffix_products(D[:, nrr:])
    s_new_r, s_new_c = lg._suffix_products(Dn[:, :nrr_out]), lg._suffix_products(Dn[:, nrr_out:])
    dig = np.zeros((nsig, ndim, 5), dtype=np.int64)
    dig[:, :, 0] = Dn
    for k, a in enumerate(axes):
        if a < nrr:
            dig[:, k, 1] = s_old_r[:, a]
        else:
            dig[:, k, 2] = s_old_c[:, a - nrr]
        if k < nrr_out:
            dig[:, k, 3] = s_new_r[:, k]
        else:
            dig[:, k, 4] = s_new_c[:, k - nrr_out]
    tasks = np.stack([D.prod(axis=1), st.w_off, st.n_in, st.n_out, st.n_out_pad, row0,
                      np.full(nsig, ndim), np.arange(nsig) * ndim], axis=1)
    return tasks, rows, dig.reshape(-1, 5), shapes


def test_native_plan_tables_match_numpy_restatement(emulated_device):
    """fs_lie_plan_tables against the numpy restatement on tensors with
    several leg counts, missing blocks, Fortran-ordered (dagger) storage and
    row/column splits, bit for bit."""
    from frostspin_amd import SU2SymmetricTensor
    from frostspin_amd import _ops

    d8 = np.array([[1, 2, 1], [1, 2, 3]])
    chi = np.array([[2, 3, 2, 1], [1, 2, 3, 4]])
    r2 = np.array([[1], [2]])
    rng = np.random.default_rng(11)
    r3 = np.array([[1], [3]])
    sparse = SU2SymmetricTensor.random((r3, r3, r3), (r3,), rng=rng)
    sparse = sparse @ sparse.dagger()  # one block among many allowed
    tensors = [
        (SU2SymmetricTensor.random((d8, chi), (d8, r2), rng=rng), [((3, 0), (2, 1)), ((1,), (0, 2, 3))]),
        (SU2SymmetricTensor.random((r2, r2, d8), (chi, d8), rng=rng), [((0, 4, 2), (3, 1)), ((4, 0), (1, 2, 3))]),
        (sparse, [((3, 2, 1), (0, 5, 4)), ((0,), (1, 2, 3, 4, 5))]),
    ]
    tensors.append((tensors[0][0].dagger(), [((0, 1), (2, 3)), ((2,), (3, 0, 1))]))
    for t, perms in tensors:
        for r_ax, c_ax in perms:
            axes = tuple(r_ax) + tuple(c_ax)
            ref = _numpy_plan_tables(t, axes, len(r_ax))
            reps = t._row_reps + t._col_reps
            st = type(t)._permute_structure(
                tuple(tuple(int(x) for x in r[1]) for r in reps), tuple(bool(s) for s in t._signature),
                t._nrr, axes, len(r_ax))
            rs_b, cs_b = t._block_strides()
            got = _ops.lie_plan_tables(
                st.pattern, np.concatenate([np.asarray(r[0], dtype=np.int64) for r in reps]),
                t._present_map(st.in_targets), t._boffs, rs_b, cs_b, nrows=int(st.n_in.sum() + st.n_out.sum()))
            for g, e in zip(got, ref, strict=True):
                assert g.shape == e.shape and np.array_equal(g, e)


@pytest.mark.gpu
def test_plan_rejects_weights_of_another_pattern(gpu):
    """fs_lie_plan_build with the shared weights of a different permute pattern
    that has the same task count and W layout must fail with FS_ERR_INVALID
    (weights carry a fingerprint of their pattern and the W layout of every
    task), not apply the other pattern's matrices."""
    import ctypes

    from frostspin_amd import SU2SymmetricTensor, _lib
    from frostspin_amd._lib import lib

    irreps = ((1, 3), (2, 4), (1, 3), (2,))
    sig = (False, True, False, True)
    # same tensor, two permutes: equal task counts and W layouts, other matrices
    st1 = SU2SymmetricTensor._permute_structure(irreps, sig, 2, (0, 1, 2, 3), 1)
    st2 = SU2SymmetricTensor._permute_structure(irreps, sig, 2, (2, 0, 3, 1), 3)
    assert len(st1.n_in) == len(st2.n_in)
    t = SU2SymmetricTensor.random(
        tuple(np.array([[1] * len(ir), list(ir)]) for ir in irreps[:2]),
        tuple(np.array([[1] * len(ir), list(ir)]) for ir in irreps[2:]),
        signature=np.array(sig), rng=np.random.default_rng(0),
    )
    degen = np.ones((sum(len(ir) for ir in irreps),), dtype=np.int64)
    in_block = np.ascontiguousarray(t._present_map(st2.in_targets), dtype=np.int64)
    rs_b, cs_b = (np.ascontiguousarray(x, dtype=np.int64) for x in t._block_strides())
    boffs = np.ascontiguousarray(t._boffs, dtype=np.int64)
    shapes = np.empty((len(st2.out_targets), 2), dtype=np.int64)
    handle = ctypes.c_void_p()
    status = lib.fs_lie_plan_build(
        ctypes.byref(st2.pattern), degen.size, _lib.i64p(degen), _lib.i64p(in_block), boffs.size,
        _lib.i64p(boffs), _lib.i64p(rs_b), _lib.i64p(cs_b), st1.weights.handle(), t._data.numel(),
        _lib.i64p(shapes), ctypes.byref(handle),
    )
    assert status == 1  # FS_ERR_INVALID
    assert handle.value is None
    assert b"weights of another pattern" in lib.fs_last_error()
    # the matching weights build a plan
    status = lib.fs_lie_plan_build(
        ctypes.byref(st2.pattern), degen.size, _lib.i64p(degen), _lib.i64p(in_block), boffs.size,
        _lib.i64p(boffs), _lib.i64p(rs_b), _lib.i64p(cs_b), st2.weights.handle(), t._data.numel(),
        _lib.i64p(shapes), ctypes.byref(handle),
    )
    assert status == 0 and handle.value is not None
    lib.fs_lie_plan_destroy(handle.value)
