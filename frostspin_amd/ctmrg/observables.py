"""Transfer-matrix spectrum of the boundary MPS (ctmrg/observables.py:4-154).

The transfer matrix acts on a left edge state by adding the cell's T1 (top)
and T3 (bottom) tensors, right to left; with A_list the site tensor and its
conjugate are sandwiched in between. Every application is a chain of device
permutes and grouped GEMMs; its leading eigenvalues come from the device
Krylov-Schur solver behind SymmetricTensor.eigs. Leg orders, the contraction
sequence and the normalization by the leading eigenvalue follow the reference
so that correlation lengths are interchangeable.
"""

import numpy as np


def compute_mps_transfer_spectrum(
    T1s, T3s, nvals, *, A_list=None, dmax_full=200, rng=None, maxiter=4000, arpack_tol=0, verbosity=0
):
    """nvals largest-magnitude eigenvalues of the transfer matrix, divided by the
    leading one (real if their imaginary parts are negligible). Returns NaNs if
    the eigensolver fails, as the reference does."""
    ST = type(T1s[0])
    Ts_up = [T.permute((3, 1, 2), (0,)) for T in T1s]
    Ts_down = [T.permute((3,), (2, 0, 1)) for T in T3s]
    if A_list is None:
        reps = (Ts_up[0].col_reps[0], Ts_down[0].col_reps[0])
        sig0 = np.array([not Ts_up[0].signature[3], not Ts_down[0].signature[1]])

        def matmat(st0):
            st = st0.permute((0,), (1, 2))
            for mu, md in zip(Ts_up, Ts_down, strict=True):
                st = (mu @ st).permute((3, 1, 2), (0, 4))
                st = (md @ st).permute((1,), (0, 2))
            return st.permute((0, 1), (2,))

    else:
        As = [A.permute((0, 1, 4, 5), (3, 2)) for A in A_list]
        Acs = [A.dagger().permute((2, 3), (4, 5, 1, 0)) for A in A_list]
        reps = (Ts_up[0].col_reps[0], As[0].col_reps[0], Acs[0].col_reps[2], Ts_down[0].col_reps[0])
        sig0 = np.array(
            [
                not Ts_up[0].signature[3],
                not As[0].signature[4],
                not Acs[0].signature[4],
                not Ts_down[0].signature[1],
            ]
        )

        def matmat(st0):
            st = st0.permute((0,), (1, 2, 3, 4))
            for mu, md, A, Ac in zip(Ts_up, Ts_down, As, Acs, strict=True):
                st = (mu @ st).permute((3, 1), (0, 2, 4, 5, 6))
                st = (A @ st).permute((0, 1, 6, 5), (2, 3, 4, 7, 8))
                st = (Ac @ st).permute((5, 2, 0), (1, 3, 4, 6))
                st = (md @ st).permute((3,), (2, 1, 0, 4))
            return st.permute((0, 1, 2, 3), (4,))

    try:
        vals = ST.eigs(
            matmat, nvals, reps, sig0, dmax_full=dmax_full, rng=rng, maxiter=maxiter,
            arpack_tol=arpack_tol, compute_vectors=False,
        ).toarray()
    except Exception as err:  # noqa: BLE001 - the reference also reports and returns NaN
        print("WARNING: transfer matrix spectrum computation failed")
        print("Error:", err)
        return np.nan * np.ones((nvals,))
    order = np.abs(vals).argsort()[: -nvals - 1 : -1]
    vals = vals[order]
    vals = vals / vals[0]
    if np.linalg.norm(np.imag(vals)) < 1e-6:
        vals = np.real(vals)
    elif verbosity > 0:
        print("Info: CTMRG transfer matrix eigenvalues are not real")
    return vals
