"""Largest-magnitude eigenpairs of an implicitly defined real operator with the
Krylov-Schur method (G. W. Stewart, "A Krylov-Schur algorithm for large
eigenproblems", SIAM J. Matrix Anal. Appl. 23 (2001) 601-614).

This is the device counterpart of the ARPACK calls of the reference's
_sparse_eig (symmetric_tensor.py:1250-1434, scipy.sparse.linalg.eigs / eigsh).
The Krylov basis is one device buffer of m+1 rows; the operator application,
the two-pass classical Gram-Schmidt (Daniel-Gragg-Kaufman-Stewart), the vector
norms and the restart rotation V <- Q^T V are device launches (grouped GEMM,
weighted norm, scale, copy). Only the projected (m+1) x m matrix lives on the
host, where its real Schur form is computed and reordered once per restart
(m is a few tens, so that work is microseconds); each Arnoldi step reads back
2(j+1)+1 numbers.

Convergence follows ARPACK's test: a Ritz pair (lambda, y) of the projected
matrix is converged when |beta * e_m^T y| <= tol * max(eps^(2/3), |lambda|),
tol = machine epsilon when 0 is passed.
"""

import numpy as np
import scipy.linalg as sla

from .. import _ops

_EPS = np.finfo(np.float64).eps


class KrylovNoConvergence(RuntimeError):
    """Raised with the converged subset when maxiter restarts were not enough."""

    def __init__(self, msg, eigenvalues, eigenvectors):
        super().__init__(msg)
        self.eigenvalues = eigenvalues
        self.eigenvectors = eigenvectors


def _gemv(V, vrow0, nrows, d, x_buf, x_off, x_stride, y_buf, y_off, y_stride, *, alpha, beta, transpose):
    """y = alpha * op(V[vrow0 : vrow0 + nrows]) x + beta * y on the device.
    Without transpose: y (nrows) = V_rows x (d); with transpose: y (d) = V_rows^T x (nrows)."""
    if transpose:
        group = (d, 1, nrows, vrow0 * d, 1, d, x_off, x_stride, 1, y_off, y_stride)
    else:
        group = (nrows, 1, d, vrow0 * d, d, 1, x_off, x_stride, 1, y_off, y_stride)
    _ops.grouped_gemm([group], V, x_buf, y_buf, alpha=alpha, beta=beta)


class _Basis:
    """m+1 device vectors of length d, stored as rows, plus the host projected matrix."""

    def __init__(self, d, m):
        self.d = d
        self.m = m
        self.V = _ops.empty((m + 1) * d)
        self.H = np.zeros((m + 1, m))
        # per-step readback: two Gram-Schmidt coefficient vectors and one norm^2
        self.coef = _ops.empty(2 * (m + 1) + 1)

    def row(self, j):
        return self.V[j * self.d : (j + 1) * self.d]

    def inject(self, j, v0):
        """V[j] <- v0 orthogonalized against V[:j] (two passes) and normalized.
        Returns False if v0 lies in span(V[:j])."""
        d, m = self.d, self.m
        w = self.row(j)
        _ops.copy2d([(1, d, 0, 1, 1, j * d, 1, 1, 0, 0, 0, 1.0, 0.0)], v0.reshape(-1), self.V)
        c = self.coef
        n0 = float(_ops.download(_ops.weighted_norm2(w, np.array([0, d]), [1.0]))[0])
        for _ in range(2 if j else 0):
            _gemv(self.V, 0, j, d, self.V, j * d, 1, c, 0, 1, alpha=1.0, beta=0.0, transpose=False)
            _gemv(self.V, 0, j, d, c, 0, 1, self.V, j * d, 1, alpha=-1.0, beta=1.0, transpose=True)
        n2 = float(_ops.download(_ops.weighted_norm2(w, np.array([0, d]), [1.0]))[0])
        if not n2 > (1e3 * _EPS) ** 2 * n0:
            return False
        _ops.scale_(w, 1.0 / np.sqrt(n2))
        return True

    def expand(self, apply, j):
        """Arnoldi step j: V[j+1] = A V[j] orthogonalized against V[0..j]. Returns
        the new subdiagonal entry (0 on an invariant subspace)."""
        d, m = self.d, self.m
        w = self.row(j + 1)
        apply(self.row(j), w)
        c = self.coef
        # pass 1: h = V w ; w -= V^T h
        _gemv(self.V, 0, j + 1, d, self.V, (j + 1) * d, 1, c, 0, 1, alpha=1.0, beta=0.0, transpose=False)
        _gemv(self.V, 0, j + 1, d, c, 0, 1, self.V, (j + 1) * d, 1, alpha=-1.0, beta=1.0, transpose=True)
        # pass 2 (re-orthogonalization): h2 = V w ; w -= V^T h2
        _gemv(self.V, 0, j + 1, d, self.V, (j + 1) * d, 1, c, m + 1, 1, alpha=1.0, beta=0.0, transpose=False)
        _gemv(self.V, 0, j + 1, d, c, m + 1, 1, self.V, (j + 1) * d, 1, alpha=-1.0, beta=1.0, transpose=True)
        n2 = _ops.weighted_norm2(w, np.array([0, d]), [1.0])
        c[2 * (m + 1) : 2 * (m + 1) + 1].copy_(n2)
        host = _ops.download(c)
        self.H[: j + 1, j] += host[: j + 1] + host[m + 1 : m + 2 + j]
        beta = float(np.sqrt(max(host[2 * (m + 1)], 0.0)))
        self.H[j + 1, j] = beta
        scale = np.abs(self.H[: j + 2, : j + 1]).max()
        if beta <= 16 * _EPS * max(scale, np.finfo(np.float64).tiny):
            self.H[j + 1, j] = 0.0
            return 0.0
        _ops.scale_(w, 1.0 / beta)
        return beta

    def rotate(self, Z, lo, n, p):
        """V[lo:p] <- Z[:, :p-lo]^T V[lo:n] (the active rows; locked rows below lo
        stay), then V[p] <- V[n] (the residual direction)."""
        d = self.d
        na, pa = n - lo, p - lo
        if pa > 0:
            Zt = _ops.upload(np.ascontiguousarray(Z[:, :pa].T))
            tmp = _ops.empty(pa * d)
            _ops.grouped_gemm([(pa, d, na, 0, na, 1, lo * d, d, 1, 0, d)], Zt, self.V, tmp)
            _ops.copy2d([(1, pa * d, 0, 1, 1, lo * d, 1, 1, 0, 0, 0, 1.0, 0.0)], tmp, self.V)
        if n != p:
            _ops.copy2d([(1, d, n * d, 1, 1, p * d, 1, 1, 0, 0, 0, 1.0, 0.0)], self.V, self.V)

    def ritz_vectors(self, Y, n):
        """Device d x k matrix (row-major) V[:n]^T Y for real Y (n x k)."""
        d = self.d
        k = Y.shape[1]
        Yd = _ops.upload(np.ascontiguousarray(Y))
        X = _ops.empty(d * k)
        _ops.grouped_gemm([(d, k, n, 0, 1, d, 0, k, 1, 0, k)], self.V, Yd, X)
        return X


def _ritz(T, b, symmetric):
    """Ritz values, vectors (of T) and residual norms |b . y| for the projected
    matrix T with residual row b."""
    if symmetric:
        vals, Y = np.linalg.eigh(0.5 * (T + T.T))
    else:
        vals, Y = np.linalg.eig(T)
    Y = Y / np.linalg.norm(Y, axis=0)
    res = np.abs(b @ Y)
    return vals, Y, res


def _blocks(S):
    """(start, size) of the diagonal blocks of a real quasi-triangular matrix."""
    out, i, n = [], 0, S.shape[0]
    while i < n:
        size = 2 if i + 1 < n and S[i + 1, i] != 0.0 else 1
        out.append((i, size))
        i += size
    return out


def _sorted_schur(T):
    """Real Schur form T = Z S Z^T with the diagonal blocks in non-increasing
    eigenvalue magnitude (selection sort by LAPACK dtrexc block moves)."""
    S, Z = sla.schur(T, output="real")
    pos = 0
    while pos < S.shape[0]:
        blks = [(i, sz) for i, sz in _blocks(S) if i >= pos]
        mags = [np.abs(np.linalg.eigvals(S[i : i + sz, i : i + sz])).max() for i, sz in blks]
        best = int(np.argmax(mags))
        if blks[best][0] != pos:
            S, Z, info = sla.lapack.dtrexc(S, Z, blks[best][0] + 1, pos + 1)
            if info != 0:  # blocks too close to swap: keep the current order
                pos += blks[0][1]
                continue
        pos += 2 if pos + 1 < S.shape[0] and S[pos + 1, pos] != 0.0 else 1
    return S, Z


def _converged(vals, res, tol, scale):
    """ARPACK's test res <= tol * max(eps^(2/3), |lambda|), with a floor of a few
    rounding units of the operator scale (the smallest residual that the
    floating-point matvec can resolve)."""
    return res <= np.maximum(tol * np.maximum(_EPS ** (2.0 / 3.0), np.abs(vals)), 8 * _EPS * scale)


def _select_largest(vals, count):
    """Indices of the `count` largest-magnitude values, completing a complex
    conjugate pair if the cut falls inside one."""
    order = np.argsort(-np.abs(vals), kind="stable")
    sel = list(order[:count])
    if count < len(vals):
        last = vals[order[count - 1]]
        if abs(last.imag) > 0 and not any(
            abs(vals[i] - np.conj(last)) <= 1e-12 * max(abs(last), 1.0) for i in sel if i != order[count - 1]
        ):
            sel.append(order[count])
    return np.array(sel, dtype=np.int64)


def largest_eigenpairs(apply, d, k, v0, *, ncv=None, maxiter=4000, tol=0.0, compute_vectors=False,
                       symmetric=False, rng=None):
    """k largest-magnitude eigenvalues (sorted by decreasing magnitude) of the
    d x d real operator `apply(x, y): y = A x` (x, y: 1-D device views of
    length d). Returns (values, vectors) where vectors is a device d x k
    row-major buffer (real eigenvalues only) or None. With symmetric the
    operator is taken as real symmetric (Lanczos case): Ritz values are real.

    Degenerate eigenvalues: a Krylov space grown from one vector holds one
    eigenvector per distinct eigenvalue. Converged leading Schur vectors are
    therefore locked (kept, residual coefficients zeroed, excluded from later
    reorderings) and, once the k wanted values are all locked, one more cycle
    is grown from a fresh random vector orthogonal to them: further copies of
    a wanted eigenvalue show up there and the iteration goes on; otherwise the
    locked values are the answer."""
    if not 0 < k < d - 1:
        raise ValueError(f"need 0 < k < d - 1 (k={k}, d={d})")
    if rng is None:
        rng = np.random.default_rng()
    m = int(ncv) if ncv is not None else min(d, max(2 * k + 1, 20))
    m = min(max(m, k + 3), d)
    max_lock = min(k + 1, m - 2)
    tol = float(tol) if tol > 0 else _EPS
    B = _Basis(d, m)
    if not B.inject(0, v0):
        raise ValueError("starting vector is zero")
    p = nlock = 0
    probed = False
    last = None
    for _ in range(int(maxiter)):
        n = m
        for j in range(p, m):
            if B.expand(apply, j) == 0.0:
                n = j + 1  # invariant subspace reached: its Ritz values are exact
                break
        T = B.H[:n, :n].copy()
        bvec = B.H[n, :n].copy()
        vals, Y, res = _ritz(T, bvec, symmetric)
        scale = np.abs(vals).max(initial=0.0)
        want = _select_largest(vals, k)
        last = (vals, Y, res, n, want)

        # restart on the active part T[nlock:, nlock:]
        na = n - nlock
        T22 = T[nlock:, nlock:]
        keep = min(k + max(1, (m - k) // 2), m - 1)
        sdim = max(1, min(keep - nlock, na - 1)) if n == m else na
        S2, Z2 = _sorted_schur(T22)
        if sdim < na and S2[sdim, sdim - 1] != 0.0:  # do not split a 2x2 block
            sdim = sdim + 1 if sdim + 1 < na else sdim - 1
        pnew = nlock + sdim
        b2 = bvec[nlock:] @ Z2[:, :sdim]
        top = T[:nlock, nlock:] @ Z2[:, :sdim]
        B.rotate(Z2, nlock, n, pnew)
        B.H[:] = 0.0
        B.H[:nlock, :nlock] = T[:nlock, :nlock]
        B.H[:nlock, nlock:pnew] = top
        B.H[nlock:pnew, nlock:pnew] = S2[:sdim, :sdim]
        B.H[pnew, nlock:pnew] = b2
        # lock the leading converged Schur vectors of the active part
        i = 0
        while i < sdim and nlock + i < max_lock:
            size = 2 if i + 1 < sdim and abs(S2[i + 1, i]) > 0 else 1
            lam = np.abs(np.linalg.eigvals(S2[i : i + size, i : i + size]))
            if not _converged(lam, np.full(size, np.linalg.norm(b2[i : i + size])), tol, scale).all():
                break
            B.H[pnew, nlock + i : nlock + i + size] = 0.0
            i += size
        if i:
            probed = False
        nlock += i
        p = pnew

        # done when the k largest are locked and nothing active can enter them
        locked = np.linalg.eigvals(B.H[:nlock, :nlock]) if nlock else np.zeros((0,))
        if locked.size >= k:
            kth = np.sort(np.abs(locked))[::-1][k - 1]
            active = np.abs(np.linalg.eigvals(B.H[nlock:pnew, nlock:pnew])) if pnew > nlock else np.zeros((0,))
            if not (active >= kth * (1.0 - 1e-10)).any() or n < m:
                if probed or nlock + 1 >= min(m, d - 1):
                    return _finish_locked(B, nlock, k, symmetric, compute_vectors)
                # grow one cycle from a fresh vector orthogonal to the locked space
                if not B.inject(nlock, _ops.upload(rng.standard_normal(d))):
                    return _finish_locked(B, nlock, k, symmetric, compute_vectors)
                B.H[nlock:, :] = 0.0
                B.H[:nlock, nlock:] = 0.0
                p = nlock
                probed = True
                continue
        elif n < m:
            # the active Krylov space closed before k values were found: restart it
            if nlock + 1 >= min(m, d - 1) or not B.inject(pnew, _ops.upload(rng.standard_normal(d))):
                return _finish_locked(B, nlock, min(k, nlock), symmetric, compute_vectors)
            B.H[pnew, :] = 0.0
    vals, Y, res, n, want = last
    ok = want[_converged(vals[want], res[want], tol, np.abs(vals).max(initial=0.0))]
    v, x = _finish(B, vals, Y, ok, n, len(ok), compute_vectors) if ok.size else (np.zeros((0,)), None)
    raise KrylovNoConvergence(f"Krylov-Schur: {ok.size}/{k} eigenvalues converged", v, x)


def _finish_locked(B, nlock, k, symmetric, compute_vectors):
    """Eigenpairs of the locked block: an invariant subspace of A up to the
    convergence tolerance."""
    T = B.H[:nlock, :nlock]
    vals, Y, _ = _ritz(T, np.zeros(nlock), symmetric)
    want = _select_largest(vals, k)
    return _finish(B, vals, Y, want, nlock, k, compute_vectors)


def _finish(B, vals, Y, want, n, k, compute_vectors):
    order = want[np.argsort(-np.abs(vals[want]), kind="stable")][:k]
    out = vals[order]
    if np.all(out.imag == 0):
        out = out.real
    if not compute_vectors:
        return out, None
    Ysel = Y[:, order]
    if np.iscomplexobj(out) or np.abs(Ysel.imag).max(initial=0.0) > 0:
        raise NotImplementedError("complex eigenvectors are not supported by the float64 device tensors")
    return out, B.ritz_vectors(np.ascontiguousarray(Ysel.real), n)
