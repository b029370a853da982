"""frostspin_amd: MI355X-native engine for frostspin's block-sparse SymmetricTensor
hot path (constructors, permute, @, truncated_svd) and the CTMRG driver on top.

Mirrors the public API of frostspin/__init__.py:1-21. Device work runs in
hand-written HIP kernels (libfrostspin_amd.so, C-ABI in include/frostspin_amd.h);
torch provides device buffers and streams only.
"""

from .symmetric_tensor import (
    AsymmetricTensor,
    DiagonalTensor,
    O2SymmetricTensor,
    SU2SymmetricTensor,
    U1SymmetricTensor,
    get_gemm_precision,
    get_symmetric_tensor_type,
    set_gemm_precision,
)
from .symmetric_tensor.tools import ASSERT_TOL

__version__ = "0.1.0"

__all__ = [
    "ASSERT_TOL",
    "AsymmetricTensor",
    "DiagonalTensor",
    "O2SymmetricTensor",
    "SU2SymmetricTensor",
    "U1SymmetricTensor",
    "__version__",
    "get_gemm_precision",
    "get_symmetric_tensor_type",
    "set_gemm_precision",
]
