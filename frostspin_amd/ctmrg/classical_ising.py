"""C4v classical CTMRG for the square-lattice Ising model with trivial or Z2
symmetry (examples/ising_classical_ctmrg.py:1-471, the reference's own driver
for the plain-method configuration: beta, chi, n_iter given).

Z2 is realised as U(1) with int8 charges {-128 (odd), 0 (even)} on bonds
rotated by K = [[-1, 1], [1, 1]] / sqrt(2) (ising_classical_ctmrg.py:358-412).
The corner is a DiagonalTensor of eigenvalues; one iteration grows the corner,
diagonalises it with truncated_eigh(chi) and renormalises the edge.
"""

import numpy as np
import scipy.linalg as lg
import scipy.special

from ..symmetric_tensor.asymmetric_tensor import AsymmetricTensor
from ..symmetric_tensor.diagonal_tensor import DiagonalTensor
from ..symmetric_tensor.u1_symmetric_tensor import U1SymmetricTensor

BETA_CRIT = 0.5 * np.log(1 + np.sqrt(2))
K_Z2 = np.array([[-1.0, 1.0], [1.0, 1.0]]) / np.sqrt(2)


def exact_magn(beta):
    """Onsager-Yang spontaneous magnetization (ising_classical_ctmrg.py:26-29)."""
    if beta < BETA_CRIT:
        return 0.0
    return (1 - np.sinh(2 * beta) ** -4) ** 0.125


def exact_energy(beta):
    """Onsager energy per site (ising_classical_ctmrg.py:32-35)."""
    k = np.sinh(2 * beta) ** -2
    x = scipy.special.ellipk(4 * k * (1 + k) ** -2)
    return -(1.0 + 2 / np.pi * (2 * np.tanh(2 * beta) ** 2 - 1) * x) / np.tanh(2 * beta)


def site_tensors(beta):
    """Dense site tensor a0 and magnetization insertion (ising_classical_ctmrg.py:39-58)."""
    h = -2 * np.eye(2) + 1
    sq_bw = lg.sqrtm(np.exp(-beta * h)).real
    delta = np.zeros((2, 2, 2, 2))
    delta[0, 0, 0, 0] = delta[1, 1, 1, 1] = 1
    a0 = np.einsum("abcd,ai,bj,ck,dl->ijkl", delta, sq_bw, sq_bw, sq_bw, sq_bw)
    m = delta.copy()
    m[1, 1, 1, 1] = -1
    magn = np.einsum("abcd,ai,bj,ck,dl->ijkl", m, sq_bw, sq_bw, sq_bw, sq_bw)
    return a0, magn


def initial_state(beta, chi, symmetry, rng, *, script_order=True):
    """Site tensors (two signatures) and random symmetric initial s, T in the
    reference's draw order (ising_classical_ctmrg.py:62-111 and 358-412). With
    script_order, a Z2 start first consumes the trivial run's draws exactly as the
    reference script does (it runs the trivial case first with the same rng)."""
    a0, magn = site_tensors(beta)
    siga = np.array([False, False, True, True])
    siga2 = np.array([False, True, True, True])
    sigT = np.array([False, False, True])
    if symmetry == "trivial":
        bond = np.array([2])
        a = AsymmetricTensor.from_array(a0, (bond,) * 2, (bond,) * 2, signature=siga)
        a2 = AsymmetricTensor._build(a.row_reps, a.col_reps, a._data, a._bshapes, a._boffs,
                                     a.block_irreps, siga2)
        s0 = rng.normal(size=(chi,))
        s0 = s0[np.abs(s0).argsort()[::-1]] / lg.norm(s0)
        crep = np.array([chi])
        s = DiagonalTensor((s0,), crep, a.block_irreps, [1], a.symmetry())
        T0 = rng.normal(size=(chi, 2, chi))
        T0 = T0 + T0.swapaxes(0, 2)
        T0 = T0 / lg.norm(T0)
        T = AsymmetricTensor.from_array(T0, (crep,), (bond, crep), signature=sigT)
        return s, T, a, a2, a0, magn
    if symmetry != "Z2":
        raise ValueError("symmetry must be 'trivial' or 'Z2'")
    if script_order:
        rng.normal(size=(chi,))
        rng.normal(size=(chi, 2, chi))
    bond = np.array([-128, 0], dtype=np.int8)
    a0d = np.einsum("abcd,ai,bj,ck,dl->ijkl", a0, K_Z2, K_Z2, K_Z2, K_Z2)
    magnd = np.einsum("abcd,ai,bj,ck,dl->ijkl", magn, K_Z2, K_Z2, K_Z2, K_Z2)
    a = U1SymmetricTensor.from_array(a0d, (bond,) * 2, (bond,) * 2, signature=siga)
    a2 = U1SymmetricTensor._build(a.row_reps, a.col_reps, a._data, a._bshapes, a._boffs,
                                  a.block_irreps, siga2)
    do = de = chi // 2
    s_odd = rng.normal(size=(do,))
    s_odd = s_odd[np.abs(s_odd).argsort()[::-1]] / lg.norm(s_odd)
    s_even = rng.normal(size=(de,))
    s_even = s_even[np.abs(s_even).argsort()[::-1]] / lg.norm(s_even)
    crep = np.array([-128] * do + [0] * de, dtype=np.int8)
    s = DiagonalTensor((s_odd, s_even), crep, np.array([-128, 0], dtype=np.int8), [1, 1], a.symmetry())
    T0 = U1SymmetricTensor.random((crep,), (bond, crep), signature=sigT, rng=rng)
    temp = U1SymmetricTensor._build(T0.row_reps, T0.col_reps, T0._data, T0._bshapes, T0._boffs,
                                    T0.block_irreps, np.array([True, False, False]))
    T = T0 + temp.permute((2,), (1, 0))
    T /= T.norm()
    return s, T, a, a2, a0d, magnd


def iterate_ctm(s, T, a, a2, chi):
    """One C4v move: grow the corner, truncated_eigh, renormalize the edge
    (ising_classical_ctmrg.py:115-204)."""
    Tp = T.permute((0, 1), (2,))
    matC = (Tp * s) @ T
    matC = a2 @ matC.permute((1, 2), (0, 3))
    matC = matC.permute((0, 2), (1, 3))
    news, U = matC.truncated_eigh(chi)
    news /= news.norm()
    newT = Tp @ U.permute((1,), (0, 2))
    newT = a @ newT.permute((1, 2), (0, 3))
    newT = U.dagger() @ newT.permute((0, 2), (1, 3))
    newT /= newT.norm()
    return news, newT


def env_1site(s, T):
    """Environment of one site (ising_classical_ctmrg.py:207-252)."""
    corner = T.permute((0, 1), (2,)) * s
    half = corner @ corner.permute((0,), (1, 2))
    full = half.permute((1, 2), (0, 3)) @ half.permute((3, 0), (1, 2))
    return full.permute((0, 3), (2, 1))


def env_2sites(s, T):
    """Environment of two neighbouring sites (ising_classical_ctmrg.py:255-304)."""
    Tp = T.permute((0, 1), (2,))
    env = Tp * s
    env = env @ env.permute((0,), (1, 2))
    env = Tp @ env.permute((0,), (1, 2, 3))
    return env.permute((1, 2, 3), (0, 4)) @ env.permute((4, 0), (1, 2, 3))


def observables(s, T, a0, magn):
    """Magnetization and energy per site from dense environments
    (ising_classical_ctmrg.py:313-351)."""
    env1 = env_1site(s, T).toarray()
    env2 = env_2sites(s, T).toarray()
    m = (env1.ravel() @ magn.ravel()) / (env1.ravel() @ a0.ravel())
    nr = np.einsum("abcdef,bgdc,afeg->", env2, a0, a0)
    sds = np.einsum("abcdef,bgdc,afeg->", env2, magn, magn) / nr
    return float(m), float(-2 * sds)


def corr_length(T):
    """Correlation length from the edge transfer matrix: with T as (chi | D, chi),
    the row transfer matrix is E[(a, c), (b, e)] = sum_d T[a, d, b] T[c, d, e]
    and xi = 1 / ln(|l0| / |l1|) from its two largest-magnitude eigenvalues.
    E is built on the device as a block matrix (a, c | b, e) from two permuted
    copies of T, and its leading eigenvalues come from the device eigensolver
    (Krylov-Schur on blocks above a hundred rows). The Z2 ({-128, 0} in int8)
    and trivial charges are self-conjugate, so the copy contracted over d may
    carry the opposite signature on that leg without changing its blocks."""
    Tl = T.permute((0, 2), (1,))  # (a, b | d)
    Tr = T.permute((1,), (0, 2))  # (d | c, e)
    sig = Tr.signature.copy()
    sig[0] = not Tl.signature[2]
    Tr = type(T)._build(Tr.row_reps, Tr.col_reps, Tr._data, Tr._bshapes, Tr._boffs, Tr.block_irreps, sig,
                        Tr._fortran, Tr.shape)
    E = (Tl @ Tr).permute((0, 2), (1, 3))  # (a, c | b, e)
    lam = np.concatenate([np.atleast_1d(np.asarray(b)) for b in E.truncated_eig(2, compute_vectors=False)
                          .diagonal_blocks])
    lam = np.sort(np.abs(lam))[::-1]
    return float(1.0 / np.log(lam[0] / lam[1]))


def run(beta=0.44, chi=32, n_iter=200, symmetry="Z2", seed=42):
    """Config driver: returns dict of observables after n_iter iterations."""
    rng = np.random.default_rng(seed)
    s, T, a, a2, a0, magn = initial_state(beta, chi, symmetry, rng)
    for _ in range(n_iter):
        s, T = iterate_ctm(s, T, a, a2, chi)
    m, e = observables(s, T, a0, magn)
    return {"magnetization": m, "energy": e, "xi": corr_length(T), "s": s, "T": T}
