"""Pinning of the O(2) oracle (CPU only): fusion parity counting, reflection
invariance and charge conservation of the dense arrays, round trips, and the
permute through U(1) against the dense route."""

import numpy as np
import pytest

from oracle import abelian
from oracle import o2 as oo2

R1 = np.array([[1, 1, 1, 1], [-1, 0, 1, 2]])
R2 = np.array([[2, 1, 1, 1], [-1, 0, 1, 2]])
R3 = np.array([[1, 2, 1, 1], [-1, 0, 1, 2]])
R4 = np.array([[1, 1, 1, 1], [-1, 0, 2, 3]])
R5 = np.array([[2, 1, 1, 1], [-1, 0, 2, 3]])


def _reflect_all(arr, reps):
    out = arr
    for k, r in enumerate(reps):
        perm, sign = oo2.reflection(r)
        out = np.moveaxis(np.moveaxis(out, k, 0)[perm] * sign.reshape((-1,) + (1,) * (out.ndim - 1)), 0, k)
    return out


def test_o2_fusion_counts():
    reps = [R1, R4, np.array([[2, 1], [1, 3]])]
    tot = oo2.combine(reps)
    sz = abelian.fuse([oo2.rep_to_u1(r) for r in reps], [False] * 3)
    counts = dict(zip(tot[1].tolist(), tot[0].tolist(), strict=True))
    for n in range(1, int(sz.max()) + 1):
        assert counts.get(n, 0) == int((sz == n).sum())
    # 0even - 0odd = trace of the reflection on the charge-0 states
    shape = [oo2.rep_dim(r) for r in reps]
    trace = 0
    for s in np.flatnonzero(sz == 0):
        idx = np.unravel_index(s, shape)
        img, sign = 0, 1
        for k, r in enumerate(reps):
            perm, sg = oo2.reflection(r)
            img = img * shape[k] + int(perm[idx[k]])
            sign *= int(sg[idx[k]])
        if img == s:
            trace += sign
    assert counts.get(0, 0) - counts.get(-1, 0) == trace
    assert counts.get(0, 0) + counts.get(-1, 0) == int((sz == 0).sum())


CASES = [((R1, R2, R3), (R4, R5), [0, 0, 0, 1, 1]), ((R2, R4), (R1,), [1, 0, 1])]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_o2_dense_invariance_and_round_trip(case):
    rows, cols, sig = CASES[case]
    sig = np.array(sig, dtype=bool)
    t = oo2.O2Tensor.random(rows, cols, sig, np.random.default_rng(case))
    arr = t.toarray()
    assert np.abs(_reflect_all(arr, rows + cols) - arr).max() < 1e-14
    u1 = abelian.Tensor.from_array(arr, [oo2.rep_to_u1(r) for r in rows], [oo2.rep_to_u1(r) for r in cols], sig)
    assert np.abs(u1.toarray() - arr).max() == 0.0
    back = oo2.O2Tensor.from_array(arr, rows, cols, sig)
    assert list(back.block_irreps) == list(t.block_irreps)
    assert max(np.abs(a - b).max() for a, b in zip(back.blocks, t.blocks, strict=True)) < 1e-14
    assert abs(t.norm() - np.linalg.norm(arr)) < 1e-13
    # random non-invariant input: projection is idempotent and invariant
    noise = np.random.default_rng(9).standard_normal(arr.shape)
    y = oo2.O2Tensor.from_array(noise, rows, cols, sig).toarray()
    assert np.abs(_reflect_all(y, rows + cols) - y).max() < 1e-13
    assert np.abs(oo2.O2Tensor.from_array(y, rows, cols, sig).toarray() - y).max() < 1e-13


@pytest.mark.parametrize("case", range(len(CASES)))
def test_o2_permute_matches_dense_route(case):
    rows, cols, sig = CASES[case]
    sig = np.array(sig, dtype=bool)
    t = oo2.O2Tensor.random(rows, cols, sig, np.random.default_rng(20 + case))
    arr = t.toarray()
    n = len(rows) + len(cols)
    axes = tuple(int(a) for a in np.random.default_rng(case).permutation(n))
    nrr = 2
    tp = t.permute(axes[:nrr], axes[nrr:])
    reps = [(rows + cols)[a] for a in axes]
    dense = oo2.O2Tensor.from_array(arr.transpose(axes), reps[:nrr], reps[nrr:], sig[list(axes)])
    assert list(tp.block_irreps) == list(dense.block_irreps)
    assert max(np.abs(a - b).max() for a, b in zip(tp.blocks, dense.blocks, strict=True)) < 1e-14
