"""Symmetry registry and norm check (symmetric_tensor/tools.py:5-39)."""

import numpy as np

# Structural/norm consistency checks run inside `assert` statements like the
# reference (disabled by python -O, README.md:46). They cost device reductions
# and host synchronisations, so they can also be switched off at run time
# through config["debug_checks"] (the package-wide configuration dict).
from ..config import ASSERT_TOL, config  # noqa: F401

symmetric_tensor_types = {}


def get_symmetric_tensor_type(symmetry):
    """SymmetricTensor subclass implementing `symmetry` ('trivial', 'U1', ...)."""
    try:
        st_type = symmetric_tensor_types[symmetry]
    except KeyError:
        msg = f"Unknown symmetry '{symmetry}'"
        raise RuntimeError(msg) from None
    return st_type


def _norm_of(t):
    import torch

    if isinstance(t, torch.Tensor):
        return float(t.double().norm())
    if hasattr(t, "norm"):
        return float(t.norm())
    return float(np.linalg.norm(t))


def check_norm(t1, t2, *, tol=ASSERT_TOL):
    """Debug check (runs under assert): both norms agree to relative tol."""
    if not config["debug_checks"]:
        return True
    n1, n2 = (_norm_of(t) for t in (t1, t2))
    dn = abs(n1 - n2)
    b = dn > tol * max(n1, n2)
    if b:
        print(f"WARNING: norm is different: {dn:.1e} > {tol * max(n1, n2):.1e}")
    return not b
