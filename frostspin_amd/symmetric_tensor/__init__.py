from .asymmetric_tensor import AsymmetricTensor
from .diagonal_tensor import DiagonalTensor
from .symmetric_tensor import get_gemm_precision, set_gemm_precision
from .tools import get_symmetric_tensor_type
from .o2_symmetric_tensor import O2SymmetricTensor
from .su2_symmetric_tensor import SU2SymmetricTensor
from .u1_symmetric_tensor import U1SymmetricTensor

__all__ = [
    "AsymmetricTensor",
    "DiagonalTensor",
    "O2SymmetricTensor",
    "SU2SymmetricTensor",
    "U1SymmetricTensor",
    "get_gemm_precision",
    "get_symmetric_tensor_type",
    "set_gemm_precision",
]
