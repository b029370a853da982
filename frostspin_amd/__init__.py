"""frostspin_amd: MI355X-native engine for frostspin's block-sparse SymmetricTensor
hot path (constructors, permute, @, truncated_svd) and the CTMRG driver on top.

Mirrors the public API of frostspin/__init__.py:1-21 (with `config`, whose
"SU2_CG_file" entry selects a Clebsch-Gordan table as in frostspin/config.py). Device work runs in
hand-written HIP kernels (libfrostspin_amd.so, C-ABI in include/frostspin_amd.h);
torch provides device buffers and streams only.
"""

from .config import ASSERT_TOL, __version__, config
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

__all__ = [
    "ASSERT_TOL",
    "AsymmetricTensor",
    "DiagonalTensor",
    "O2SymmetricTensor",
    "SU2SymmetricTensor",
    "U1SymmetricTensor",
    "__version__",
    "config",
    "get_gemm_precision",
    "get_symmetric_tensor_type",
    "set_gemm_precision",
]
