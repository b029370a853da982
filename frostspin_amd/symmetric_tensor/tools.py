"""Symmetry registry and norm check (symmetric_tensor/tools.py:5-39)."""

import numpy as np

ASSERT_TOL = 4e-13

# Structural/norm consistency checks run inside `assert` statements like the
# reference (disabled by python -O, README.md:46). They cost device reductions
# and host synchronisations, so they can also be switched off at run time.
config = {"debug_checks": __debug__}

symmetric_tensor_types = {}


def get_symmetric_tensor_type(symmetry):
    """SymmetricTensor subclass implementing `symmetry` ('trivial', 'U1', ...)."""
    try:
        st_type = symmetric_tensor_types[symmetry]
    except KeyError:
        msg = f"Unknown symmetry '{symmetry}'"
        raise RuntimeError(msg) from None
    return st_type


def check_norm(t1, t2, *, tol=ASSERT_TOL):
    """Debug check (runs under assert): both norms agree to relative tol."""
    if not config["debug_checks"]:
        return True
    n1 = t1.norm() if hasattr(t1, "norm") else float(np.linalg.norm(t1))
    n2 = t2.norm() if hasattr(t2, "norm") else float(np.linalg.norm(t2))
    dn = abs(n1 - n2)
    b = dn > tol * max(n1, n2)
    if b:
        print(f"WARNING: norm is different: {dn:.1e} > {tol * max(n1, n2):.1e}")
    return not b
