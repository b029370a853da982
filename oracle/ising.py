"""Oracle: C4v classical Ising CTMRG with Z2 (int8 {-128, 0}) or trivial
(all-zero charges) symmetry, numpy blocks. Restates
examples/ising_classical_ctmrg.py:23-351 on oracle.abelian.Tensor.

Test infrastructure only (the product never imports it).

Pinned by (tests/test_oracle.py): Onsager's closed-form energy (and
magnetization in the ordered phase) far from criticality, where chi = 16
converges to ~1e-8; Z2 and trivial runs agree; the edge transfer-matrix
correlation length (this build's own observable: the reference defines none)
approaches Onsager's exact axis value xi^-1 = -ln tanh(beta) - 2 beta from
below (1.4 % short at beta = 0.3, chi = 16) and grows with chi near the
critical point. The reference's own example run at beta = 0.44, chi = 32, 200
iterations (tests/golden/ref_ising.npz) pins the whole trajectory.
"""

import numpy as np
import scipy.linalg as lg
import scipy.special

from .abelian import Tensor

BETA_CRIT = 0.5 * np.log(1 + np.sqrt(2))
K = np.array([[-1.0, 1.0], [1.0, 1.0]]) / np.sqrt(2)


def exact_energy(beta):
    """(ising_classical_ctmrg.py:32-35)"""
    k = np.sinh(2 * beta) ** -2
    x = scipy.special.ellipk(4 * k * (1 + k) ** -2)
    return -(1.0 + 2 / np.pi * (2 * np.tanh(2 * beta) ** 2 - 1) * x) / np.tanh(2 * beta)


def exact_magn(beta):
    """(ising_classical_ctmrg.py:26-29)"""
    return 0.0 if beta < BETA_CRIT else (1 - np.sinh(2 * beta) ** -4) ** 0.125


def exact_xi(beta):
    """Axis correlation length of the disordered phase: 1/xi = -ln tanh(beta) - 2 beta."""
    return 1.0 / (-np.log(np.tanh(beta)) - 2 * beta)


def _site(beta):
    h = -2 * np.eye(2) + 1
    sq = lg.sqrtm(np.exp(-beta * h)).real
    delta = np.zeros((2, 2, 2, 2))
    delta[0, 0, 0, 0] = delta[1, 1, 1, 1] = 1
    a0 = np.einsum("abcd,ai,bj,ck,dl->ijkl", delta, sq, sq, sq, sq)
    m = delta.copy()
    m[1, 1, 1, 1] = -1
    magn = np.einsum("abcd,ai,bj,ck,dl->ijkl", m, sq, sq, sq, sq)
    return a0, magn


class Diag:
    def __init__(self, blocks, irreps):
        self.blocks = [np.asarray(b, dtype=np.float64) for b in blocks]
        self.irreps = np.asarray(irreps, dtype=np.int8)

    def norm(self):
        return float(np.sqrt(sum(np.sum(b * b) for b in self.blocks)))

    def scaled(self, x):
        return Diag([b * x for b in self.blocks], self.irreps)


def initial_state(beta, chi, symmetry, rng):
    a0, magn = _site(beta)
    siga = np.array([False, False, True, True])
    siga2 = np.array([False, True, True, True])
    sigT = np.array([False, False, True])
    # the reference script draws the trivial start first with the same rng
    s0 = rng.normal(size=(chi,))
    T0 = rng.normal(size=(chi, 2, chi))
    if symmetry == "trivial":
        bond = np.zeros((2,), dtype=np.int8)
        crep = np.zeros((chi,), dtype=np.int8)
        a = Tensor.from_array(a0, (bond,) * 2, (bond,) * 2, siga)
        a2 = Tensor(a.row_reps, a.col_reps, a.blocks, a.block_irreps, siga2)
        s0 = s0[np.abs(s0).argsort()[::-1]] / lg.norm(s0)
        T0 = T0 + T0.swapaxes(0, 2)
        T0 = T0 / lg.norm(T0)
        T = Tensor.from_array(T0, (crep,), (bond, crep), sigT)
        return Diag([s0], [0]), T, a, a2, a0, magn
    bond = np.array([-128, 0], dtype=np.int8)
    a0d = np.einsum("abcd,ai,bj,ck,dl->ijkl", a0, K, K, K, K)
    magnd = np.einsum("abcd,ai,bj,ck,dl->ijkl", magn, K, K, K, K)
    a = Tensor.from_array(a0d, (bond,) * 2, (bond,) * 2, siga)
    a2 = Tensor(a.row_reps, a.col_reps, a.blocks, a.block_irreps, siga2)
    h = chi // 2
    so = rng.normal(size=(h,))
    so = so[np.abs(so).argsort()[::-1]] / lg.norm(so)
    se = rng.normal(size=(h,))
    se = se[np.abs(se).argsort()[::-1]] / lg.norm(se)
    crep = np.array([-128] * h + [0] * h, dtype=np.int8)
    T0 = Tensor.random((crep,), (bond, crep), sigT, rng)
    temp = Tensor(T0.row_reps, T0.col_reps, T0.blocks, T0.block_irreps, [True, False, False])
    T = T0 + temp.permute((2,), (1, 0))
    T = T * (1.0 / T.norm())
    return Diag([so, se], [-128, 0]), T, a, a2, a0d, magnd


def iterate_ctm(s, T, a, a2, chi):
    """(ising_classical_ctmrg.py:115-204)"""
    Tp = T.permute((0, 1), (2,))
    matC = Tp.diag_mul(s.blocks, s.irreps) @ T
    matC = a2 @ matC.permute((1, 2), (0, 3))
    matC = matC.permute((0, 2), (1, 3))
    vals, irr, U = matC.truncated_eigh(chi)
    news = Diag(vals, irr)
    news = news.scaled(1.0 / news.norm())
    newT = Tp @ U.permute((1,), (0, 2))
    newT = a @ newT.permute((1, 2), (0, 3))
    newT = U.dagger() @ newT.permute((0, 2), (1, 3))
    newT = newT * (1.0 / newT.norm())
    return news, newT


def observables(s, T, a0, magn):
    """(ising_classical_ctmrg.py:207-351)"""
    Tp = T.permute((0, 1), (2,))
    corner = Tp.diag_mul(s.blocks, s.irreps)
    half = corner @ corner.permute((0,), (1, 2))
    full = half.permute((1, 2), (0, 3)) @ half.permute((3, 0), (1, 2))
    env1 = full.permute((0, 3), (2, 1)).toarray()
    env = corner @ corner.permute((0,), (1, 2))
    env = Tp @ env.permute((0,), (1, 2, 3))
    env2 = (env.permute((1, 2, 3), (0, 4)) @ env.permute((4, 0), (1, 2, 3))).toarray()
    m = (env1.ravel() @ magn.ravel()) / (env1.ravel() @ a0.ravel())
    nr = np.einsum("abcdef,bgdc,afeg->", env2, a0, a0)
    sds = np.einsum("abcdef,bgdc,afeg->", env2, magn, magn) / nr
    return float(m), float(-2 * sds)


def corr_length(T):
    return corr_length_dense(T.toarray())


def corr_length_dense(t):
    """Row transfer matrix of the (chi, D, chi) edge tensor t:
    E[(a, c), (b, e)] = sum_d t[a, d, b] t[c, d, e]; xi = 1 / ln(|l0| / |l1|)
    from LAPACK eigvals of the dense (chi^2 x chi^2) matrix."""
    chi = t.shape[0]
    E = np.einsum("adb,cde->acbe", t, t).reshape(chi * chi, chi * chi)
    ev = np.linalg.eigvals(E)
    ev = ev[np.abs(ev).argsort()[::-1]]
    return float(1.0 / np.log(np.abs(ev[0]) / np.abs(ev[1])))


def run(beta=0.44, chi=32, n_iter=200, symmetry="Z2", seed=42):
    rng = np.random.default_rng(seed)
    s, T, a, a2, a0, magn = initial_state(beta, chi, symmetry, rng)
    for _ in range(n_iter):
        s, T = iterate_ctm(s, T, a, a2, chi)
    m, e = observables(s, T, a0, magn)
    return {"magnetization": m, "energy": e, "xi": corr_length(T), "s": s, "T": T}
