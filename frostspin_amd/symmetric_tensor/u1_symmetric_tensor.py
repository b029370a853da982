"""U(1) symmetric tensors (u1_symmetric_tensor.py:22-57); Z2 is U(1) with int8
charges {0, -128} since -128 + -128 wraps to 0."""

import numpy as np

from .. import _lib
from .abelian_symmetric_tensor import AbelianSymmetricTensor
from .tools import symmetric_tensor_types


class U1SymmetricTensor(AbelianSymmetricTensor):
    """SymmetricTensor with global U(1) symmetry."""

    _symmetry = "U1"

    @staticmethod
    def combine_representations(reps, signature):
        """Row-major Kronecker sum of (+/-) charges, int8 wrap-around (native planner)."""
        signature = np.asarray(signature, dtype=bool)
        assert signature.shape == (len(reps),)
        return _lib.u1_fuse(list(reps), signature.astype(np.uint8))

    @staticmethod
    def conjugate_irrep(irr):
        if np.ndim(irr) == 0:
            return np.int8((-int(irr) + 128) % 256 - 128)  # int8 wrap-around (-(-128) = -128)
        return U1SymmetricTensor.conjugate_representation(irr)

    @staticmethod
    def conjugate_representation(rep):
        return np.negative(np.asarray(rep, dtype=np.int8), dtype=np.int8)

    def toU1(self):
        return self


symmetric_tensor_types["U1"] = U1SymmetricTensor
