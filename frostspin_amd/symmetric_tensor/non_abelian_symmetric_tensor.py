"""Non-abelian block tensors: representations are int arrays (2, n), row 0 the
degeneracies and row 1 the irreps (non_abelian_symmetric_tensor.py:7-90)."""

import numpy as np

from .symmetric_tensor import SymmetricTensor


class NonAbelianSymmetricTensor(SymmetricTensor):
    @staticmethod
    def init_representation(degen, irreps):
        rep = np.empty((2, len(degen)), dtype=np.int64)
        rep[0] = degen
        rep[1] = irreps
        return rep

    @classmethod
    def get_block_sizes(cls, row_reps, col_reps, signature):
        """Irreps shared by the fused rows and columns and the block shapes
        (non_abelian_symmetric_tensor.py:36-67)."""
        nrr = len(row_reps)
        signature = np.asarray(signature, dtype=bool)
        row_tot = cls.combine_representations(tuple(row_reps), signature[:nrr])
        col_tot = cls.combine_representations(tuple(col_reps), ~signature[nrr:])
        rinds, cinds = (row_tot[1][:, None] == col_tot[1]).nonzero()
        block_irreps = np.ascontiguousarray(row_tot[1][rinds], dtype=np.int64)
        block_shapes = np.empty((rinds.size, 2), dtype=np.int64)
        block_shapes[:, 0] = row_tot[0][rinds]
        block_shapes[:, 1] = col_tot[0][cinds]
        return block_irreps, block_shapes

    def check_blocks_fit_representations(self):
        assert self._block_irreps.size == self._nblocks
        if self._nblocks == 0:
            return True
        row_rep = self.get_row_representation()
        col_rep = self.get_column_representation()
        ri = row_rep[1].searchsorted(self._block_irreps)
        ci = col_rep[1].searchsorted(self._block_irreps)
        assert (ri < row_rep.shape[1]).all() and (ci < col_rep.shape[1]).all()
        assert (row_rep[1, ri] == self._block_irreps).all()
        assert (col_rep[1, ci] == self._block_irreps).all()
        for b in range(self._nblocks):
            assert tuple(self._bshapes[b]) == (row_rep[0, ri[b]], col_rep[0, ci[b]])
        return True

    def totrivial(self):
        from .asymmetric_tensor import AsymmetricTensor

        ar = self.toarray()
        rr = tuple(np.array([d]) for d in self._shape[: self._nrr])
        cr = tuple(np.array([d]) for d in self._shape[self._nrr :])
        return AsymmetricTensor.from_array(ar, rr, cr, signature=self.signature)
