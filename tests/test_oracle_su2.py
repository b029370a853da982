"""Pinning of the SU(2) oracle (CPU only).

* Clebsch-Gordan: the reference's own coefficient table (fixture from its data
  file) and unitarity of the coupled basis;
* fusion counts: Sz-weight counting;
* from_array / toarray: the reference's known-answer blocks
  (test_su2_tensor.py:24-30, 55-68), round trips, SU(2) invariance of the dense
  tensors (real y-rotations on every leg plus signed Sz conservation), the
  projector property on non-invariant input;
* permute: the dense-route identity of test_su2_tensor.py:36-79 and norm
  preservation.
"""

import os

import numpy as np
import pytest
import scipy.linalg as lg

from oracle import abelian
from oracle import su2 as osu2

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "su2_cg_reference.npz")


def reference_cg_table():
    table = {}
    with np.load(GOLDEN) as data:
        for key in data.files:
            if key[0] != "i":
                continue
            sh = tuple(int(x) for x in key[2:].split("_"))
            flat = np.zeros((int(np.prod(sh)),))
            flat[data[key]] = data["c" + key[1:]]
            table[sh] = flat.reshape(sh)
    return table


def test_oracle_cg_matches_reference_table():
    table = reference_cg_table()
    assert len(table) == 91
    for sh, ref in table.items():
        assert np.abs(osu2.clebsch_gordan(*sh) - ref).max() < 2e-16, sh


@pytest.mark.parametrize(("d1", "d2"), [(2, 2), (3, 2), (4, 5), (6, 6), (7, 4)])
def test_oracle_cg_coupled_basis_is_orthogonal(d1, d2):
    cols = [osu2.clebsch_gordan(d1, d2, d3).reshape(d1 * d2, d3) for d3 in range(abs(d1 - d2) + 1, d1 + d2, 2)]
    U = np.concatenate(cols, axis=1)
    assert U.shape == (d1 * d2, d1 * d2)
    assert np.abs(U.T @ U - np.eye(d1 * d2)).max() < 1e-14


def _sz_weights(rep):
    """Multiset of 2*Sz over a representation."""
    out = []
    for d, irr in zip(rep[0], rep[1], strict=True):
        out += list(np.arange(irr - 1, -irr - 1, -2)) * int(d)
    return np.array(out)


def test_oracle_combine_counts_multiplets():
    reps = [np.array([[1, 2, 1], [1, 2, 3]]), np.array([[2, 1], [2, 4]]), np.array([[1], [3]])]
    tot = osu2.combine(reps)
    w = _sz_weights(reps[0])
    for r in reps[1:]:
        w = (w[:, None] + _sz_weights(r)[None, :]).reshape(-1)
    assert int(tot[0] @ tot[1]) == w.size
    for d, irr in zip(tot[0], tot[1], strict=True):
        # multiplicity of spin (irr-1)/2 = N(2Sz = irr-1) - N(2Sz = irr+1)
        assert d == (w == irr - 1).sum() - (w == irr + 1).sum()


SDS_22 = np.array(
    [[0.25, 0.0, 0.0, 0.0], [0.0, -0.25, 0.5, 0.0], [0.0, 0.5, -0.25, 0.0], [0.0, 0.0, 0.0, 0.25]]
)
SDS_22B = np.array(
    [[-0.25, 0.0, 0.0, -0.5], [0.0, 0.25, 0.0, 0.0], [0.0, 0.0, 0.25, 0.0], [-0.5, 0.0, 0.0, -0.25]]
)
R = np.array([[1], [2]])


def test_oracle_known_answer_blocks():
    for arr, sig in ((SDS_22, [0, 0, 1, 1]), (SDS_22B, [0, 1, 1, 0])):
        a = arr.reshape(2, 2, 2, 2)
        st = osu2.SU2Tensor.from_array(a, (R, R), (R, R), sig)
        assert list(st.block_irreps) == [1, 3]
        assert abs(st.blocks[0][0, 0] + 0.75) < 1e-14
        assert abs(st.blocks[1][0, 0] - 0.25) < 1e-14
        assert np.abs(st.toarray() - a).max() < 1e-14
        assert abs(st.norm() - lg.norm(arr)) < 1e-14


def _spin_y_rotation(d, theta):
    s = (d - 1) / 2
    m = s - np.arange(d)
    sp = np.zeros((d, d))
    for i in range(1, d):
        sp[i - 1, i] = np.sqrt(s * (s + 1) - m[i] * (m[i] + 1))  # |m> -> |m+1>
    return lg.expm(-theta * (sp - sp.T) / 2)  # exp(-i theta S_y), real


def _rotate_all_legs(arr, reps, theta):
    out = arr
    for k, r in enumerate(reps):
        blocks = [_spin_y_rotation(irr, theta) for d, irr in zip(r[0], r[1], strict=True) for _ in range(d)]
        g = lg.block_diag(*blocks)
        out = np.moveaxis(np.tensordot(g, out, axes=(1, k)), 0, k)
    return out


CASES = [
    ((np.array([[1, 2, 1], [1, 2, 3]]), np.array([[2, 1], [2, 4]])),
     (np.array([[1, 2, 1], [1, 2, 3]]), R), [0, 1, 1, 0]),
    ((R, R, np.array([[1, 1], [1, 3]])), (np.array([[2, 1], [1, 3]]),), [1, 0, 0, 1]),
    ((np.array([[2, 1], [1, 3]]),), (R, np.array([[1, 1], [2, 4]])), [0, 0, 1]),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_oracle_dense_tensors_are_su2_invariant(case):
    rows, cols, sig = CASES[case]
    sig = np.array(sig, dtype=bool)
    t = osu2.SU2Tensor.random(rows, cols, sig, np.random.default_rng(case))
    arr = t.toarray()
    reps = rows + cols
    for theta in (0.3, 1.7):
        assert np.abs(_rotate_all_legs(arr, reps, theta) - arr).max() < 1e-13
    # signed Sz conservation: the cast to U(1) keeps every coefficient
    charges = [_sz_weights(r).astype(np.int8) for r in reps]
    u1 = abelian.Tensor.from_array(arr, charges[: len(rows)], charges[len(rows) :], sig)
    assert np.abs(u1.toarray() - arr).max() == 0.0
    # round trip and projector property
    t2 = osu2.SU2Tensor.from_array(arr, rows, cols, sig)
    assert max(np.abs(a - b).max() for a, b in zip(t.blocks, t2.blocks, strict=True)) < 1e-14
    noise = np.random.default_rng(7).standard_normal(arr.shape)
    y = osu2.SU2Tensor.from_array(noise, rows, cols, sig).toarray()
    assert np.abs(_rotate_all_legs(y, reps, 0.9) - y).max() < 1e-12
    y2 = osu2.SU2Tensor.from_array(y, rows, cols, sig).toarray()
    assert np.abs(y2 - y).max() < 1e-13


@pytest.mark.parametrize("case", range(len(CASES)))
def test_oracle_permute_preserves_norm_and_inverts(case):
    rows, cols, sig = CASES[case]
    sig = np.array(sig, dtype=bool)
    t = osu2.SU2Tensor.random(rows, cols, sig, np.random.default_rng(10 + case))
    n = t.ndim
    axes = tuple(np.random.default_rng(case).permutation(n))
    tp = t.permute(axes[:2], axes[2:])
    assert abs(tp.norm() - t.norm()) < 1e-13
    inv = tuple(np.argsort(axes))
    back = tp.permute(inv[: t.nrr], inv[t.nrr :])
    assert max(np.abs(a - b).max() for a, b in zip(back.blocks, t.blocks, strict=True)) < 1e-13
