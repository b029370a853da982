"""DiagonalTensor: singular values / eigenvalues, kept on the host.

Diagonal weights are tiny (at most chi values) and are needed on the host anyway
for the truncation decision, so they stay numpy arrays as in the reference
(frostspin/symmetric_tensor/diagonal_tensor.py:5-276). A device copy is made on
demand when they scale a SymmetricTensor.
"""

import numpy as np

from .. import _ops


class DiagonalTensor:
    """
    Diagonal tensor as produced by SVD or eigenvalue computation.
    Interface follows diagonal_tensor.py:10-276.
    """

    def __init__(self, diag_blocks, representation, block_irreps, block_degen, symmetry):
        self._diagonal_blocks = tuple(np.asarray(db) for db in diag_blocks)
        self._representation = representation
        self._block_irreps = np.asarray(block_irreps)
        self._block_degen = np.asarray(block_degen)
        self._symmetry = symmetry
        self._nblocks = len(diag_blocks)
        self._dev = None
        assert self._block_degen.shape == (self._nblocks,)
        assert self._block_irreps.shape[0] == self._nblocks
        assert all(db.ndim == 1 for db in self._diagonal_blocks)

    # getters -------------------------------------------------------------
    @property
    def diagonal_blocks(self):
        return self._diagonal_blocks

    @property
    def block_degen(self):
        return self._block_degen

    @property
    def block_irreps(self):
        return self._block_irreps

    @property
    def nblocks(self):
        return self._nblocks

    @property
    def ndim(self):
        return 1

    @property
    def representation(self):
        return self._representation

    @property
    def signature(self):
        return np.array([False, True])

    @property
    def shape(self):
        return (len(self),)

    @property
    def dtype(self):
        return self._diagonal_blocks[0].dtype

    def _device_values(self):
        """Concatenated values on the device and per-block offsets (cached)."""
        if self._dev is None:
            offs = np.zeros((self._nblocks + 1,), dtype=np.int64)
            offs[1:] = np.cumsum([db.size for db in self._diagonal_blocks])
            cat = (
                np.concatenate([np.asarray(db, dtype=np.float64) for db in self._diagonal_blocks])
                if self._nblocks
                else np.zeros((0,))
            )
            self._dev = (_ops.upload(cat), offs)
        return self._dev

    def _new(self, blocks):
        return type(self)(
            blocks, self._representation, self._block_irreps, self._block_degen, self._symmetry
        )

    # magic methods -------------------------------------------------------
    def __repr__(self):
        return f"DiagonalTensor with {self.nblocks} blocks and {self.symmetry()} symmetry"

    def __len__(self):
        return int(
            sum(self._diagonal_blocks[i].size * self._block_degen[i] for i in range(self._nblocks))
        )

    def __mul__(self, x):
        if isinstance(x, DiagonalTensor):
            if not (self._representation == x.representation).all():
                raise ValueError("Representations do not match")
            blocks = tuple(b1 * b2 for b1, b2 in zip(self._diagonal_blocks, x.diagonal_blocks, strict=True))
        elif np.issubdtype(type(x), np.number):
            blocks = tuple(x * db for db in self._diagonal_blocks)
        else:
            return NotImplemented
        return self._new(blocks)

    def __rmul__(self, x):
        if np.issubdtype(type(x), np.number):
            return self * x
        msg = f"Unsupported operation for type: {type(x)}"
        raise TypeError(msg)

    def __imul__(self, x):
        if not np.issubdtype(type(x), np.number):
            raise NotImplementedError("Invalid type for *=")
        for db in self._diagonal_blocks:
            db[:] *= x
        self._dev = None
        return self

    def __truediv__(self, x):
        return self * (1.0 / x)

    def __rtruediv__(self, x):
        if np.issubdtype(type(x), np.number):
            return self._new(tuple(x / db for db in self._diagonal_blocks))
        raise NotImplementedError("Invalid type for /")

    def __itruediv__(self, x):
        self *= 1.0 / x
        return self

    def __pow__(self, x):
        if not np.issubdtype(type(x), np.number):
            raise NotImplementedError("Invalid type for **")
        return self._new(tuple(db**x for db in self._diagonal_blocks))

    def __neg__(self):
        return self._new(tuple(-db for db in self._diagonal_blocks))

    def __add__(self, other):
        if not isinstance(other, DiagonalTensor):
            raise NotImplementedError("Invalid type for +")
        if not (self._representation == other.representation).all():
            raise ValueError("Representations do not match")
        return self._new(
            tuple(b1 + b2 for b1, b2 in zip(self._diagonal_blocks, other.diagonal_blocks, strict=True))
        )

    def __sub__(self, other):
        return self + (-other)

    def __pos__(self):
        return self

    # misc ----------------------------------------------------------------
    def copy(self):
        return self._new(tuple(db.copy() for db in self._diagonal_blocks))

    def symmetry(self):
        return self._symmetry

    def sum(self):
        s = 0.0
        for i in range(self._nblocks):
            s += self._block_degen[i] * self._diagonal_blocks[i].sum()
        return s

    def norm(self):
        n2 = 0.0
        for i in range(self._nblocks):
            n2 += self._block_degen[i] * float(np.sum(np.abs(self._diagonal_blocks[i]) ** 2))
        return np.sqrt(n2)

    def toarray(self, *, sort=False):
        """Weights grouped by irrep, or sorted by non-increasing magnitude if sort."""
        dt = self.dtype if self._nblocks else np.float64
        arr = np.empty(self.shape, dtype=dt)
        k = 0
        for i in range(self._nblocks):
            deg = self._block_degen[i]
            dim = self._diagonal_blocks[i].size
            arr[k : k + deg * dim].reshape(dim, deg)[:] = self._diagonal_blocks[i][:, None]
            k += deg * dim
        assert k == arr.size
        if sort:
            arr = arr[np.abs(arr).argsort()[::-1]]
        return arr

    # I/O (npz schema of diagonal_tensor.py:238-276) -----------------------
    def save_to_file(self, savefile):
        np.savez_compressed(savefile, **self.get_data_dic())

    def get_data_dic(self, *, prefix=""):
        data = {
            prefix + "_block_irreps": self._block_irreps,
            prefix + "_block_degen": self._block_degen,
            prefix + "_representation": self._representation,
            prefix + "_symmetry": self._symmetry,
        }
        for bi, db in enumerate(self._diagonal_blocks):
            data[f"{prefix}_diagonal_block_{bi}"] = db
        return data

    @classmethod
    def load_from_dic(cls, data, *, prefix=""):
        symmetry = str(data[prefix + "_symmetry"])
        block_irreps = data[prefix + "_block_irreps"]
        block_degen = data[prefix + "_block_degen"]
        representation = data[prefix + "_representation"]
        diag_blocks = tuple(data[f"{prefix}_diagonal_block_{bi}"] for bi in range(block_irreps.size))
        return cls(diag_blocks, representation, block_irreps, block_degen, symmetry)

    @classmethod
    def load_from_file(cls, savefile, *, prefix=""):
        with np.load(savefile) as fin:
            return cls.load_from_dic(fin, prefix=prefix)
