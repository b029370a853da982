"""Tensor without symmetry (asymmetric_tensor.py:7-105): one dense block.

A representation is a 1-element integer array holding the leg dimension. Data
movement reuses the charge-gather engine with every charge equal to zero,
which makes the whole matrix a single block.
"""

import numpy as np

from .. import _ops
from .abelian_symmetric_tensor import ChargeGatherEngine
from .symmetric_tensor import SymmetricTensor
from .tools import symmetric_tensor_types


class AsymmetricTensor(ChargeGatherEngine, SymmetricTensor):
    """Tensor with no symmetry, mostly for debug and benchmarks purposes."""

    _symmetry = "trivial"
    _irrep = np.zeros([1], dtype=np.int8)

    @staticmethod
    def singlet():
        return np.ones((1,), dtype=int)

    @staticmethod
    def combine_representations(reps, _signature):
        return np.prod(reps, axis=0)

    @staticmethod
    def conjugate_irrep(irr):
        return irr

    @staticmethod
    def conjugate_representation(rep):
        return rep

    @staticmethod
    def init_representation(degen, _irreps):
        return np.asarray(degen).reshape((1,))

    @staticmethod
    def representation_dimension(rep):
        return int(np.asarray(rep).reshape(-1)[0])

    @staticmethod
    def irrep_dimension(_rep):
        return 1

    @classmethod
    def _leg_charges(cls, rep):
        return np.zeros((cls.representation_dimension(rep),), dtype=np.int8)

    @classmethod
    def get_block_sizes(cls, row_reps, col_reps, _signature):
        nr = int(np.prod([cls.representation_dimension(r) for r in row_reps]))
        nc = int(np.prod([cls.representation_dimension(r) for r in col_reps]))
        return cls._irrep.copy(), np.array([[nr, nc]], dtype=np.int64)

    @classmethod
    def _blocks_from_dense(cls, arr, row_reps, col_reps, _signature):
        nr = int(np.prod([cls.representation_dimension(r) for r in row_reps]))
        nc = int(np.prod([cls.representation_dimension(r) for r in col_reps]))
        data = _ops.upload(np.asarray(arr, dtype=np.float64))
        return data, np.array([[nr, nc]], dtype=np.int64), cls._irrep.copy()

    def _tomatrix(self):
        m, n = (int(x) for x in self._bshapes[0])
        flat = _ops.download(self._data[int(self._boffs[0]) : int(self._boffs[0]) + m * n])
        return flat.reshape(n, m).T.copy() if self._fortran else flat.reshape(m, n)

    def _transpose_order(self):
        return np.arange(self._nblocks), self._block_irreps.copy()

    def dual(self):
        return type(self)._build(
            self._row_reps, self._col_reps, self._data, self._bshapes, self._boffs,
            self._block_irreps, ~self._signature, self._fortran,
        )

    def check_blocks_fit_representations(self):
        if self._nblocks > 0:
            assert self._nblocks == 1
            assert self._block_irreps.shape == (1,)
            assert self._block_irreps[0] == 0
            assert tuple(self._bshapes[0]) == self.matrix_shape
        return True

    def totrivial(self):
        return self

    def update_signature(self, sign_update):
        up = np.asarray(sign_update, dtype=bool)
        assert up.shape == (self._ndim,)
        self._signature = self._signature ^ up
        assert self.check_blocks_fit_representations()


symmetric_tensor_types["trivial"] = AsymmetricTensor
