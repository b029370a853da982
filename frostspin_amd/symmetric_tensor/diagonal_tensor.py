"""DiagonalTensor: the weights (singular values, eigenvalues) of a block-diagonal
matrix, one value vector per irrep block, each value standing for a multiplet
of ``block_degen`` equal entries.

Public interface and npz key schema follow
frostspin/symmetric_tensor/diagonal_tensor.py:10-276 (constructor arguments,
getters, arithmetic, ``toarray``, ``get_data_dic`` / ``load_from_dic``).

Layout: all blocks live in ONE packed host vector (``_packed``) with block
boundaries in ``_bounds``; ``diagonal_blocks`` are views into it. Weights are
at most chi values and the truncation that produces and consumes them is a
host decision, so the host copy is the primary one; the device copy that
scales a SymmetricTensor (its columns or rows) is made once, on first use, and
dropped when the values change through ``*=`` or ``/=`` (writes through the
``diagonal_blocks`` views are not tracked). Every arithmetic operation is one
vectorized expression over the packed vector.
"""

import numpy as np

from .. import _ops


def _same_array(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and bool((a == b).all())


def _is_scalar(x):
    return np.isscalar(x) and np.issubdtype(type(x), np.number)


class DiagonalTensor:
    """Block-diagonal weights with irrep multiplicities (see module docstring)."""

    def __init__(self, diag_blocks, representation, block_irreps, block_degen, symmetry):
        parts = [np.asarray(b) for b in diag_blocks]
        if not all(p.ndim == 1 for p in parts):
            raise ValueError("diagonal blocks must be one-dimensional")
        sizes = np.array([p.size for p in parts], dtype=np.int64)
        bounds = np.zeros((len(parts) + 1,), dtype=np.int64)
        np.cumsum(sizes, out=bounds[1:])
        dtype = np.result_type(*parts) if parts else np.float64
        packed = np.empty((int(bounds[-1]),), dtype=dtype)
        for p, lo, hi in zip(parts, bounds[:-1], bounds[1:], strict=True):
            packed[lo:hi] = p
        self._init_packed(packed, bounds, representation, block_irreps, block_degen, symmetry)

    def _init_packed(self, packed, bounds, representation, block_irreps, block_degen, symmetry):
        self._packed = packed
        self._bounds = bounds
        self._representation = representation
        self._block_irreps = np.asarray(block_irreps)
        self._block_degen = np.asarray(block_degen)
        self._symmetry = symmetry
        self._dev = None
        n = bounds.size - 1
        if self._block_degen.shape != (n,) or self._block_irreps.shape[0] != n:
            raise ValueError("one irrep and one degeneracy per diagonal block are required")

    def _like(self, packed):
        """Same structure, other values (a new packed vector)."""
        out = DiagonalTensor.__new__(type(self))
        out._init_packed(packed, self._bounds, self._representation, self._block_irreps, self._block_degen,
                         self._symmetry)
        return out

    def _multiplicities(self):
        """Multiplicity of every packed value (its block's degeneracy)."""
        return np.repeat(self._block_degen, np.diff(self._bounds))

    def _check_same_structure(self, other):
        if not _same_array(self._representation, other.representation):
            raise ValueError("DiagonalTensor representations differ")

    def _device_values(self):
        """All values as one float64 device vector and each block's start offset
        (made on first use; in-place changes drop it)."""
        if self._dev is None:
            self._dev = (_ops.upload(np.ascontiguousarray(self._packed, dtype=np.float64)), self._bounds)
        return self._dev

    # getters ------------------------------------------------------------------
    @property
    def diagonal_blocks(self):
        b = self._bounds
        return tuple(self._packed[b[i] : b[i + 1]] for i in range(b.size - 1))

    @property
    def block_degen(self):
        return self._block_degen

    @property
    def block_irreps(self):
        return self._block_irreps

    @property
    def nblocks(self):
        return self._bounds.size - 1

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
        return self._packed.dtype

    def symmetry(self):
        return self._symmetry

    def __len__(self):
        return int(np.dot(np.diff(self._bounds), self._block_degen)) if self.nblocks else 0

    def __repr__(self):
        return f"DiagonalTensor ({self._symmetry}) with {self.nblocks} blocks, {len(self)} values"

    # arithmetic -----------------------------------------------------------------
    def __mul__(self, x):
        if isinstance(x, DiagonalTensor):
            self._check_same_structure(x)
            return self._like(self._packed * x._packed)
        if _is_scalar(x):
            return self._like(self._packed * x)
        return NotImplemented

    def __rmul__(self, x):
        if _is_scalar(x):
            return self._like(x * self._packed)
        raise TypeError(f"cannot multiply a DiagonalTensor by {type(x)}")

    def __imul__(self, x):
        if not _is_scalar(x):
            raise NotImplementedError(f"in-place product with {type(x)}")
        self._packed *= x
        self._dev = None
        return self

    def __truediv__(self, x):
        return self * (1.0 / x)

    def __rtruediv__(self, x):
        if _is_scalar(x):
            return self._like(x / self._packed)
        raise NotImplementedError(f"cannot divide {type(x)} by a DiagonalTensor")

    def __itruediv__(self, x):
        return self.__imul__(1.0 / x)

    def __pow__(self, x):
        if not _is_scalar(x):
            raise NotImplementedError(f"power of a DiagonalTensor by {type(x)}")
        return self._like(self._packed**x)

    def __neg__(self):
        return self._like(-self._packed)

    def __pos__(self):
        return self

    def __add__(self, other):
        if not isinstance(other, DiagonalTensor):
            raise NotImplementedError(f"cannot add {type(other)} to a DiagonalTensor")
        self._check_same_structure(other)
        return self._like(self._packed + other._packed)

    def __sub__(self, other):
        return self + (-other)

    # reductions and export --------------------------------------------------------
    def copy(self):
        return self._like(self._packed.copy())

    def sum(self):
        """Sum of all diagonal entries, each value counted with its multiplicity."""
        return self._packed @ self._multiplicities() if self._packed.size else 0.0

    def norm(self):
        """Frobenius norm of the full diagonal (multiplicities included)."""
        a = np.abs(self._packed)
        return float(np.sqrt((a * a) @ self._multiplicities())) if a.size else 0.0

    def toarray(self, *, sort=False):
        """Every diagonal entry, block by block, each value repeated block_degen
        times in a row; sorted by decreasing magnitude if sort."""
        if not self.nblocks:
            return np.zeros((0,))
        arr = np.repeat(self._packed, self._multiplicities())
        if sort:
            arr = arr[np.argsort(np.abs(arr))[::-1]]
        return arr

    # npz I/O (key schema of diagonal_tensor.py:238-276) -----------------------------
    def save_to_file(self, savefile):
        np.savez_compressed(savefile, **self.get_data_dic())

    def get_data_dic(self, *, prefix=""):
        data = {
            f"{prefix}_symmetry": self._symmetry,
            f"{prefix}_representation": self._representation,
            f"{prefix}_block_irreps": self._block_irreps,
            f"{prefix}_block_degen": self._block_degen,
        }
        data.update({f"{prefix}_diagonal_block_{i}": b for i, b in enumerate(self.diagonal_blocks)})
        return data

    @classmethod
    def load_from_dic(cls, data, *, prefix=""):
        irreps = data[f"{prefix}_block_irreps"]
        blocks = [data[f"{prefix}_diagonal_block_{i}"] for i in range(irreps.size)]
        return cls(blocks, data[f"{prefix}_representation"], irreps, data[f"{prefix}_block_degen"],
                   str(data[f"{prefix}_symmetry"]))

    @classmethod
    def load_from_file(cls, savefile, *, prefix=""):
        with np.load(savefile) as fin:
            return cls.load_from_dic(fin, prefix=prefix)
