"""Pin the CPU oracle (and the native host planner) with independent checks:
closed forms, brute force on tiny inputs, dense linear algebra, Onsager."""

import itertools

import numpy as np
import pytest
import scipy.linalg as lg

from frostspin_amd import _lib
from oracle import ising
from oracle.abelian import Tensor, find_block_cuts, fuse


def brute_fuse(reps, sig):
    out = []
    for combo in itertools.product(*[list(r) for r in reps]):
        tot = sum(-int(c) if s else int(c) for c, s in zip(combo, sig, strict=True))
        out.append(((tot + 128) % 256) - 128)
    return np.array(out, dtype=np.int8)


def test_fuse_closed_forms():
    assert (fuse([np.int8([1, -1]), np.int8([0, 2])], [False, True]) == [1, -1, -1, -3]).all()
    z2 = np.int8([-128, 0])
    assert (fuse([z2, z2], [False, False]) == [0, -128, -128, 0]).all()
    assert (fuse([z2, z2], [True, False]) == [0, -128, -128, 0]).all()
    rng = np.random.default_rng(0)
    for _ in range(20):
        reps = [rng.integers(-128, 128, size=rng.integers(1, 5)).astype(np.int8) for _ in range(3)]
        sig = rng.integers(0, 2, size=3).astype(bool)
        ref = brute_fuse(reps, sig)
        assert (fuse(reps, sig) == ref).all()
        assert (_lib.u1_fuse(reps, sig.astype(np.uint8)) == ref).all()


@pytest.mark.parametrize(
    ("vals", "cut", "dims", "kw", "expect"),
    [
        ([[3.0, 2.0, 1.0], [2.5, 0.5]], 3, None, {}, [2, 1]),
        ([[1.0, 0.99999, 0.5], [0.999995]], 1, None, {"degen_ratio": 0.9999}, [2, 1]),
        ([[1.0, 1e-20]], 5, None, {}, [1, 0][:1]),
        ([[1.0], [0.9]], 3, [3, 1], {}, [1, 0]),
        ([[1.0], [0.9]], 4, [3, 1], {}, [1, 1]),
        ([[0.0]], 3, None, {}, [0]),
        ([[1.0, 0.5, 0.1], [0.2]], 10, None, {"atol": 0.15}, [2, 1]),
        ([[1.0, 0.5, 0.1], [0.2]], 10, None, {"rtol": 0.3}, [2, 0]),
    ],
)
def test_find_block_cuts_hand_cases(vals, cut, dims, kw, expect):
    vals = [np.array(v) for v in vals]
    got = find_block_cuts(vals, cut, dims=dims, **kw)
    assert list(got) == list(expect)
    # native planner implementation of the same rule
    d = np.ones(len(vals), dtype=np.int64) if dims is None else np.array(dims)
    rtol = kw.get("rtol", sum(v.size for v in vals) * np.finfo(float).eps)
    cutoff = kw.get("atol", 0.0) + rtol * max(v[0] for v in vals)
    nat = _lib.find_block_cuts_native(vals, cut, d, cutoff, kw.get("degen_ratio", 1.0))
    assert list(nat) == list(expect)


def test_find_block_cuts_native_random():
    rng = np.random.default_rng(1)
    for _ in range(200):
        nb = int(rng.integers(1, 6))
        vals = [np.sort(rng.random(rng.integers(1, 8)) ** 3)[::-1] for _ in range(nb)]
        if rng.random() < 0.5:  # planted near-degenerate pairs
            vals[0] = np.repeat(vals[0], 2) * (1 - 1e-6 * rng.random(2 * vals[0].size))
            vals[0] = np.sort(vals[0])[::-1]
        dims = rng.integers(1, 4, size=nb)
        cut = int(rng.integers(1, 15))
        dr = float(rng.choice([1.0, 0.9999, 0.99]))
        ref = find_block_cuts(vals, cut, dims=dims, degen_ratio=dr)
        rtol = sum(v.size for v in vals) * np.finfo(float).eps
        nat = _lib.find_block_cuts_native(vals, cut, dims, rtol * max(v[0] for v in vals), dr)
        assert (ref == nat).all()


def _brute_blocks(arr, row_reps, col_reps, sig):
    nrr = len(row_reps)
    rc = brute_fuse(row_reps, sig[:nrr])
    cc = brute_fuse(col_reps, ~sig[nrr:])
    m = arr.reshape(rc.size, cc.size)
    out = {}
    for irr in sorted(set(rc.tolist()) & set(cc.tolist())):
        out[irr] = np.array([[m[i, j] for j in range(cc.size) if cc[j] == irr]
                             for i in range(rc.size) if rc[i] == irr])
    return out


def test_from_array_block_order_and_roundtrip():
    rng = np.random.default_rng(2)
    for _ in range(5):
        reps = [rng.integers(-2, 3, size=rng.integers(1, 4)).astype(np.int8) for _ in range(4)]
        sig = rng.integers(0, 2, size=4).astype(bool)
        arr = rng.standard_normal(tuple(r.size for r in reps))
        charges = brute_fuse(reps, sig)
        arr[charges.reshape(arr.shape) != 0] = 0.0
        t = Tensor.from_array(arr, reps[:2], reps[2:], sig)
        ref = _brute_blocks(arr, reps[:2], reps[2:], sig)
        assert list(t.block_irreps) == list(ref)
        for irr, b in zip(t.block_irreps, t.blocks, strict=True):
            assert np.array_equal(b, ref[int(irr)])
        assert np.array_equal(t.toarray(), arr)


def test_permute_matmul_vs_dense():
    rng = np.random.default_rng(3)
    legs = [rng.integers(-2, 3, size=3).astype(np.int8) for _ in range(5)]
    sig = rng.integers(0, 2, size=5).astype(bool)
    t = Tensor.random(legs[:2], legs[2:], sig, rng)
    d = t.toarray()
    for _ in range(6):
        p = rng.permutation(5)
        k = int(rng.integers(1, 5))
        assert np.array_equal(t.permute(tuple(p[:k]), tuple(p[k:])).toarray(), d.transpose(p))
    assert np.array_equal(t.transpose().toarray(), d.transpose(2, 3, 4, 0, 1))
    b = Tensor.random(legs[2:], legs[:1], np.concatenate([~sig[2:], [True]]), rng)
    c = t @ b
    assert np.allclose(c.toarray(as_matrix=True), t.toarray(as_matrix=True) @ b.toarray(as_matrix=True),
                       atol=1e-14)


def test_truncated_svd_eigh_vs_dense():
    rng = np.random.default_rng(4)
    legs = [rng.integers(-1, 2, size=4).astype(np.int8) for _ in range(4)]
    t = Tensor.random(legs[:2], legs[2:], np.array([0, 0, 1, 1], bool), rng)
    ref = lg.svd(t.toarray(as_matrix=True), compute_uv=False)
    U, s, irr, V = t.truncated_svd(7)
    got = np.sort(np.concatenate(s))[::-1]
    assert np.abs(got - ref[: got.size]).max() < 1e-13
    rep = rng.integers(-2, 3, size=8).astype(np.int8)
    a = Tensor.random((rep,), (rep,), np.array([False, True]), rng)
    h = a @ a.dagger()
    vals, _, _ = h.truncated_eigh(5)
    got = np.sort(np.concatenate(vals))[::-1]
    ev = np.sort(lg.eigvalsh(h.toarray()))[::-1]
    assert np.abs(got - ev[: got.size]).max() < 1e-13


def test_ising_oracle_onsager():
    r = ising.run(beta=0.3, chi=16, n_iter=60, symmetry="Z2")
    assert abs(r["energy"] - ising.exact_energy(0.3)) < 1e-12
    assert abs(r["magnetization"]) < 1e-12
    rt = ising.run(beta=0.3, chi=16, n_iter=60, symmetry="trivial")
    assert abs(rt["energy"] - r["energy"]) < 1e-12
    ro = ising.run(beta=0.5, chi=16, n_iter=80, symmetry="trivial")
    assert abs(abs(ro["magnetization"]) - ising.exact_magn(0.5)) < 1e-6
    assert abs(ro["energy"] - ising.exact_energy(0.5)) < 1e-8


def test_ising_oracle_correlation_length():
    """Edge transfer-matrix correlation length against Onsager's exact axis
    value xi^-1 = -ln tanh(beta) - 2 beta (disordered phase): at beta = 0.3 it
    approaches the exact 1.5789 from below as chi grows (1.5324 at chi = 8,
    1.5565 at chi = 16). Near the critical point it grows with chi and stays
    below the exact value (about 364 at beta = 0.44)."""
    exact = ising.exact_xi(0.3)
    xi8 = ising.run(beta=0.3, chi=8, n_iter=60, symmetry="Z2")["xi"]
    xi16 = ising.run(beta=0.3, chi=16, n_iter=60, symmetry="Z2")["xi"]
    assert xi8 < xi16 < exact
    assert abs(xi16 - exact) < 0.02 * exact
    xi_c16 = ising.run(beta=0.44, chi=16, n_iter=200, symmetry="Z2")["xi"]
    xi_c32 = ising.run(beta=0.44, chi=32, n_iter=200, symmetry="Z2")["xi"]
    assert 10.0 < xi_c16 < xi_c32 < ising.exact_xi(0.44)
