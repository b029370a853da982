"""Abelian (U(1), and Z2 through int8 wrap-around) block-sparse tensors.

Interface of frostspin/symmetric_tensor/abelian_symmetric_tensor.py:253-430.
Representations are int8 charge vectors; a block exists for every charge
present in both the fused rows and the fused (conjugated) columns. All data
movement (dense <-> blocks, permute) runs through one native gather kernel
driven by planner tables (fs_abelian_permute_plan_create).
"""

import numpy as np
import torch

from .. import _ops
from .symmetric_tensor import SymmetricTensor


def _dense_source(arr):
    """Flat float64 device copy of a dense array (device arrays are used as is)."""
    if isinstance(arr, torch.Tensor):
        return arr.reshape(-1).to(dtype=_ops.DTYPE).contiguous()
    return _ops.upload(np.asarray(arr, dtype=np.float64))


class ChargeGatherEngine:
    """Data movement of charge-labelled (abelian or trivial) block tensors through
    the native planned gather: permute, dense -> blocks, blocks -> dense."""

    def _all_leg_charges(self):
        return [self._leg_charges(r) for r in self._row_reps + self._col_reps]

    def _plan(self, axes, nrr, *, new_dense=False):
        return _ops.abelian_permute_plan(
            self._shape,
            self._all_leg_charges(),
            self._signature,
            self._nrr,
            np.asarray(self._block_irreps, dtype=np.int8),
            self._boffs,
            self._fortran,
            axes,
            nrr,
            new_dense=new_dense,
        )

    @classmethod
    def _blocks_from_dense(cls, arr, row_reps, col_reps, signature):
        """Dense -> blocks: gather with the old side declared dense (abelian:363-373;
        in-block row order = stable order of matching rows as _numba_reduce_to_blocks)."""
        reps = tuple(row_reps) + tuple(col_reps)
        nrr = len(row_reps)
        plan = _ops.abelian_permute_plan(
            [cls.representation_dimension(r) for r in reps],
            [cls._leg_charges(r) for r in reps],
            signature,
            nrr,
            np.zeros((0,), np.int8),
            None,
            False,
            np.arange(len(reps)),
            nrr,
            old_dense=True,
        )
        src = _dense_source(arr)
        data = _ops.permute(plan, src)
        return data, plan.shapes, cls._irreps_dtype(plan.irreps)

    @classmethod
    def _irreps_dtype(cls, irreps):
        return np.asarray(irreps, dtype=np.int8)

    def _tomatrix(self):
        """Blocks -> dense (abelian:375-383), zeros outside allowed blocks."""
        plan = self._plan(np.arange(self._ndim), self._nrr, new_dense=True)
        dense = _ops.permute(plan, self._data)
        return _ops.download(dense).reshape(self.matrix_shape)

    def _transpose_order(self):
        # transpose maps irrep -> conjugate irrep; reorder blocks to stay sorted
        conj_irreps = self.conjugate_representation(self._block_irreps)
        so = np.argsort(conj_irreps, kind="stable")
        return so, conj_irreps[so]

    def _permute_data(self, axes, nrr):
        plan = self._plan(axes, nrr)
        data = _ops.permute(plan, self._data)
        return data, plan.shapes, self._irreps_dtype(plan.irreps)


class AbelianSymmetricTensor(ChargeGatherEngine, SymmetricTensor):
    """Tensor with an abelian symmetry; irreps are int8, reps are int8 1-D arrays."""

    @staticmethod
    def singlet():
        return np.zeros((1,), dtype=np.int8)

    @staticmethod
    def init_representation(degen, irreps):
        return np.repeat(np.asarray(irreps, dtype=np.int8), np.asarray(degen, dtype=np.int64))

    @staticmethod
    def representation_dimension(rep):
        return int(np.asarray(rep).size)

    @staticmethod
    def irrep_dimension(_rep):
        return 1

    @classmethod
    def _leg_charges(cls, rep):
        return np.asarray(rep, dtype=np.int8)

    @classmethod
    def get_block_sizes(cls, row_reps, col_reps, signature):
        """Irreps present on both sides and block shapes (abelian:287-318)."""
        row_tot = cls.combine_representations(row_reps, signature[: len(row_reps)])
        col_tot = cls.combine_representations(col_reps, ~signature[len(row_reps) :])
        row_irreps, row_sizes = np.unique(row_tot, return_counts=True)
        col_irreps, col_sizes = np.unique(col_tot, return_counts=True)
        rinds, cinds = (row_irreps[:, None] == col_irreps).nonzero()
        block_irreps = np.ascontiguousarray(row_irreps[rinds])
        block_shapes = np.array([row_sizes[rinds], col_sizes[cinds]]).T.copy()
        return block_irreps, block_shapes.astype(np.int64).reshape(-1, 2)

    def toabelian(self):
        return self

    def totrivial(self):
        from .asymmetric_tensor import AsymmetricTensor

        ar = self.toarray()
        rr = tuple(np.array([d]) for d in self._shape[: self._nrr])
        cr = tuple(np.array([d]) for d in self._shape[self._nrr :])
        return AsymmetricTensor.from_array(ar, rr, cr, signature=self.signature)

    def update_signature(self, sign_update):
        """Flip signatures; abelian charges are conjugated, coefficients unchanged."""
        up = np.asarray(sign_update, dtype=bool)
        assert up.shape == (self._ndim,)
        row_reps = list(self._row_reps)
        col_reps = list(self._col_reps)
        for i in up.nonzero()[0]:
            if i < self._nrr:
                row_reps[i] = self.conjugate_representation(row_reps[i])
            else:
                j = i - self._nrr
                col_reps[j] = self.conjugate_representation(col_reps[j])
        self._row_reps = tuple(row_reps)
        self._col_reps = tuple(col_reps)
        self._signature = self._signature ^ up
        assert self.check_blocks_fit_representations()

    def check_blocks_fit_representations(self):
        assert self._block_irreps.size == self._nblocks
        if self._nblocks == 0:
            return True
        row_irreps = self.get_row_representation()
        col_irreps = self.get_column_representation()
        for irr, (m, n) in zip(self._block_irreps, self._bshapes, strict=True):
            nr = int((row_irreps == irr).sum())
            nc = int((col_irreps == irr).sum())
            assert nr > 0
            assert nc > 0
            assert (m, n) == (nr, nc)
        return True

