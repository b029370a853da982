"""Device-resident block-sparse SymmetricTensor.

Public interface follows frostspin/symmetric_tensor/symmetric_tensor.py:35-1477
(constructors, permute, @, truncated_svd, eigh, norm, I/O...), so code written
against the reference runs unchanged. Storage is MI355X-first:

* all blocks of a tensor live in ONE 1-D float64 device buffer (``_data``);
* block b is the (rows_b x cols_b) matrix at ``_boffs[b]``; every block of a
  tensor has the same memory order (row-major, or column-major when
  ``_fortran`` is set); block offsets are a permutation of the packed prefix
  sums, so transpose() and dagger() are zero-copy relabellings;
* ``blocks`` exposes torch views of that buffer with the logical shapes.

Heavy work runs in native kernels (frostspin_amd._ops): permute via a planned
gather, ``@`` as one grouped MFMA GEMM over all matching blocks, SVD/eigh as
batched one-sided Jacobi. Host code only does bookkeeping on small integer
metadata and the truncation decision.
"""

import threading
from collections import OrderedDict

import numpy as np
import torch

from .. import _ops
from ..misc_tools.svd_tools import find_block_cuts
from .diagonal_tensor import DiagonalTensor
from .tools import check_norm
from .tools import config as _debug_config


def validated_signature(signature0, n_row_reps, n_col_reps):
    """Default signature: False on rows, True on columns (symmetric_tensor.py:24-32)."""
    if signature0 is None:
        signature = np.zeros((n_row_reps + n_col_reps,), dtype=bool)
        signature[n_row_reps:] = True
    else:
        signature = np.ascontiguousarray(signature0, dtype=bool)
        if signature.shape != (n_row_reps + n_col_reps,):
            raise ValueError("Invalid signature shape")
    return signature


def _packed_offsets(shapes):
    sizes = shapes[:, 0] * shapes[:, 1] if len(shapes) else np.zeros((0,), dtype=np.int64)
    offs = np.zeros((len(shapes),), dtype=np.int64)
    if len(shapes) > 1:
        offs[1:] = np.cumsum(sizes[:-1])
    return offs, int(sizes.sum())


def _pack_host_or_device(blocks):
    """Blocks (numpy arrays or torch tensors) -> (device buffer, shapes)."""
    blocks = list(blocks)
    shapes = np.array([tuple(b.shape) for b in blocks], dtype=np.int64).reshape(-1, 2)
    if not blocks:
        return _ops.empty(0), shapes
    if all(isinstance(b, torch.Tensor) for b in blocks):
        for b in blocks:
            if b.is_complex():
                raise NotImplementedError("complex coefficients are not supported on device")
        flat = [b.detach().reshape(-1).to(device=_ops.device(), dtype=_ops.DTYPE) for b in blocks]
        return torch.cat(flat), shapes
    arrs = []
    for b in blocks:
        a = b.detach().cpu().numpy() if isinstance(b, torch.Tensor) else np.asarray(b)
        if np.iscomplexobj(a):
            if np.abs(a.imag).max(initial=0.0) != 0.0:
                raise NotImplementedError("complex coefficients are not supported on device")
            a = a.real
        arrs.append(np.ascontiguousarray(a, dtype=np.float64).reshape(-1))
    return _ops.upload(np.concatenate(arrs)), shapes


class SymmetricTensor:
    """
    Generic base class for block-sparse symmetric tensors. Tensors are seen as
    matrices with legs grouped into rows and columns; one dense block per irrep of
    the fused row representation, blocks sorted by irrep. Some allowed blocks may be
    missing (they are zero), as produced by matmul.
    """

    _symmetry = NotImplemented

    @classmethod
    def symmetry(cls):
        return cls._symmetry

    # symmetry implementation (defined by subclasses) -----------------------
    @staticmethod
    def singlet():
        raise NotImplementedError("Must be defined in derived class")

    @staticmethod
    def combine_representations(reps, signature):
        raise NotImplementedError("Must be defined in derived class")

    @staticmethod
    def conjugate_irrep(irr):
        raise NotImplementedError("Must be defined in derived class")

    @staticmethod
    def conjugate_representation(rep):
        raise NotImplementedError("Must be defined in derived class")

    @staticmethod
    def init_representation(degen, irreps):
        raise NotImplementedError("Must be defined in derived class")

    @staticmethod
    def representation_dimension(rep):
        raise NotImplementedError("Must be defined in derived class")

    @staticmethod
    def irrep_dimension(rep):
        raise NotImplementedError("Must be defined in derived class")

    @classmethod
    def get_block_sizes(cls, row_reps, col_reps, signature):
        raise NotImplementedError("Must be defined in derived class")

    def check_blocks_fit_representations(self):
        raise NotImplementedError("Must be defined in derived class")

    def toabelian(self):
        raise NotImplementedError("Must be defined in derived class")

    def update_signature(self, sign_update):
        raise NotImplementedError("Must be defined in derived class")

    # construction ----------------------------------------------------------
    def __init__(self, row_reps, col_reps, blocks, block_irreps, signature):
        data, shapes = _pack_host_or_device(blocks)
        offs, _ = _packed_offsets(shapes)
        self._setup(row_reps, col_reps, data, shapes, offs, block_irreps, signature, False)

    def _setup(self, row_reps, col_reps, data, shapes, offs, block_irreps, signature, fortran, shape=None):
        self._row_reps = tuple(row_reps)
        self._col_reps = tuple(col_reps)
        self._shape = shape if shape is not None else tuple(
            self.representation_dimension(r) for r in self._row_reps + self._col_reps
        )
        self._ndim = len(self._shape)
        self._nrr = len(self._row_reps)
        self._signature = np.ascontiguousarray(signature, dtype=bool)
        self._block_irreps = np.asarray(block_irreps)
        self._data = data
        self._bshapes = np.asarray(shapes, dtype=np.int64).reshape(-1, 2)
        self._boffs = np.asarray(offs, dtype=np.int64).reshape(-1)
        self._fortran = bool(fortran)
        self._nblocks = self._bshapes.shape[0]
        assert 0 < self._nrr < self._ndim
        assert self._signature.shape == (self._ndim,)
        assert self._block_irreps.size == self._nblocks
        if _debug_config["debug_checks"]:
            assert sorted(set(self._block_irreps.tolist())) == list(self._block_irreps)
            assert self.check_blocks_fit_representations()

    @classmethod
    def _build(cls, row_reps, col_reps, data, shapes, offs, block_irreps, signature, fortran=False, shape=None):
        """Tensor around an existing buffer; shape (leg dimensions) may be passed
        when the caller knows it, to skip recomputing it from the representations."""
        st = cls.__new__(cls)
        st._setup(row_reps, col_reps, data, shapes, offs, block_irreps, signature, fortran, shape)
        return st

    def _with(self, data=None, *, fortran=None):
        """Same structure (reps, blocks, signature) with another buffer."""
        return type(self)._build(
            self._row_reps,
            self._col_reps,
            self._data if data is None else data,
            self._bshapes,
            self._boffs,
            self._block_irreps,
            self._signature,
            self._fortran if fortran is None else fortran,
            self._shape,
        )

    @classmethod
    def from_array(cls, arr, row_reps, col_reps, signature=None):
        """Dense array -> symmetric blocks (symmetric_tensor.py:118-146)."""
        row_reps = tuple(row_reps)
        col_reps = tuple(col_reps)
        in_reps = row_reps + col_reps
        if isinstance(arr, torch.Tensor):
            # device arrays stay on the device; host torch tensors become numpy
            arr = arr.detach()
            if arr.device != _ops.device():
                arr = arr.cpu().numpy()
        else:
            arr = np.asarray(arr)
        if tuple(arr.shape) != tuple(cls.representation_dimension(rep) for rep in in_reps):
            raise ValueError("Representations do not match array shape")
        signature = validated_signature(signature, len(row_reps), len(col_reps))
        data, shapes, block_irreps = cls._blocks_from_dense(arr, row_reps, col_reps, signature)
        offs, _ = _packed_offsets(shapes)
        st = cls._build(row_reps, col_reps, data, shapes, offs, block_irreps, signature)
        assert check_norm(st, arr)
        return st

    def toarray(self, *, as_matrix=False):
        """Dense host array with tensor shape (or matrix shape)."""
        if self._nblocks == 0:
            mat = np.zeros(self.matrix_shape, dtype=np.float64)
        else:
            mat = self._tomatrix()
        assert check_norm(self, mat)
        if as_matrix:
            return mat
        return mat.reshape(self._shape)

    @classmethod
    def from_diagonal_tensor(cls, dt):
        if type(dt) is not DiagonalTensor:
            raise ValueError(f"Invalid input type: {type(dt)}")
        if cls.symmetry() != dt.symmetry():
            raise ValueError("Symmetries do not match")
        blocks = tuple(np.diag(db) for db in dt.diagonal_blocks)
        return cls(
            (dt.representation,), (dt.representation,), blocks, dt.block_irreps,
            np.array([False, True]),
        )

    def cast(self, symmetry):
        return getattr(self, f"to{symmetry}")()

    @classmethod
    def random(cls, row_reps, col_reps, signature=None, *, rng=None):
        """Random blocks uniform in [-0.5, 0.5), drawn per block in block order with
        the host generator exactly as symmetric_tensor.py:514-546 (same seed, same
        tensor)."""
        signature = validated_signature(signature, len(row_reps), len(col_reps))
        if rng is None:
            rng = np.random.default_rng()
        block_irreps, block_shapes = cls.get_block_sizes(row_reps, col_reps, signature)
        blocks = [rng.random(tuple(sh)) - 0.5 for sh in block_shapes]
        return cls(row_reps, col_reps, blocks, block_irreps, signature)

    def copy(self):
        return self._with(self._data.clone())

    # getters ---------------------------------------------------------------
    @property
    def nblocks(self):
        return self._nblocks

    @property
    def ndim(self):
        return self._ndim

    @property
    def n_row_reps(self):
        return self._nrr

    @property
    def shape(self):
        return self._shape

    @property
    def matrix_shape(self):
        return (
            int(np.prod(self._shape[: self._nrr])),
            int(np.prod(self._shape[self._nrr :])),
        )

    @property
    def dtype(self):
        return np.dtype(np.float64)

    @property
    def ncoeff(self):
        return int((self._bshapes[:, 0] * self._bshapes[:, 1]).sum())

    @property
    def blocks(self):
        """Host numpy arrays of the blocks (symmetric_tensor.py:320 returns
        numpy arrays): one device-to-host copy of the whole buffer, so writing
        into them does not change the tensor. Device views: device_blocks."""
        if self.nblocks == 0:
            return ()
        host = _ops.download(self._data)
        if not self._data.is_cuda:
            host = host.copy()  # a host buffer would otherwise be shared
        out = []
        for (m, n), off in zip(self._bshapes, self._boffs, strict=True):
            seg = host[int(off) : int(off) + int(m * n)]
            out.append(seg.reshape(int(n), int(m)).T if self._fortran else seg.reshape(int(m), int(n)))
        return tuple(out)

    @property
    def device_blocks(self):
        """Torch views (device) of each block with its logical shape; writes go
        to the tensor's own buffer."""
        out = []
        for (m, n), off in zip(self._bshapes, self._boffs, strict=True):
            seg = self._data[int(off) : int(off) + int(m * n)]
            out.append(seg.view(int(n), int(m)).T if self._fortran else seg.view(int(m), int(n)))
        return tuple(out)

    @property
    def block_irreps(self):
        return self._block_irreps

    @property
    def row_reps(self):
        return self._row_reps

    @property
    def col_reps(self):
        return self._col_reps

    @property
    def signature(self):
        return self._signature

    def _strides(self, b):
        m, n = self._bshapes[b]
        return (1, int(m)) if self._fortran else (int(n), 1)

    def _dims(self):
        return np.array([self.irrep_dimension(r) for r in self._block_irreps], dtype=np.float64)

    # magic methods -----------------------------------------------------------
    def __repr__(self):
        bipartition = (self._nrr, self._ndim - self._nrr)
        return f"{self.symmetry()} SymmetricTensor with bipartition {bipartition} and shape {self._shape}"

    def _affine(self, alpha):
        out = _ops.empty(self._data.numel())
        _ops.axpby_(alpha, self._data, 0.0, out)
        return self._with(out)

    def __add__(self, other):
        """Blockwise sum, missing blocks on either side kept (symmetric_tensor.py:349-372)."""
        assert self.match_representations(other)
        irreps = sorted(set(self._block_irreps.tolist()) | set(other.block_irreps.tolist()))
        irreps = np.array(irreps, dtype=self._block_irreps.dtype)
        ia = {int(r): i for i, r in enumerate(self._block_irreps)}
        ib = {int(r): i for i, r in enumerate(other.block_irreps)}
        shapes = np.empty((len(irreps), 2), dtype=np.int64)
        for k, r in enumerate(irreps):
            shapes[k] = (self._bshapes[ia[int(r)]] if int(r) in ia else other._bshapes[ib[int(r)]])
        offs, total = _packed_offsets(shapes)
        out = _ops.empty(total)
        d1, d2 = [], []
        for k, r in enumerate(irreps):
            m, n = shapes[k]
            if int(r) in ia:
                b = ia[int(r)]
                rs, cs = self._strides(b)
                d1.append((m, n, self._boffs[b], rs, cs, offs[k], n, 1, 0, 0, 0, 1.0, 0.0))
            if int(r) in ib:
                b = ib[int(r)]
                rs, cs = other._strides(b)
                beta = 1.0 if int(r) in ia else 0.0
                d2.append((m, n, other._boffs[b], rs, cs, offs[k], n, 1, 0, 0, 0, 1.0, beta))
        _ops.copy2d(d1, self._data, out)
        _ops.copy2d(d2, other._data, out)
        return type(self)._build(
            self._row_reps, self._col_reps, out, shapes, offs, irreps, self._signature, shape=self._shape
        )

    def __sub__(self, other):
        return self + (-other)

    def __mul__(self, other):
        if np.issubdtype(type(other), np.number):
            return self._affine(other)
        if isinstance(other, DiagonalTensor):
            assert self._symmetry == other.symmetry()
            assert not _debug_config["debug_checks"] or (self._col_reps[-1] == other.representation).all()
            return self._diagonal_mul(other, left=False)
        raise TypeError("unsupported operation")

    def __rmul__(self, other):
        if np.issubdtype(type(other), np.number):
            return self._affine(other)
        if isinstance(other, DiagonalTensor):
            assert self._symmetry == other.symmetry()
            assert not _debug_config["debug_checks"] or (self._row_reps[0] == other.representation).all()
            return self._diagonal_mul(other, left=True)
        raise TypeError("unsupported operation")

    def __truediv__(self, x):
        return self * (1.0 / x)

    def __rtruediv__(self, x):
        return self * (1.0 / x)

    def __pos__(self):
        return self

    def __neg__(self):
        return self._affine(-1.0)

    def __imul__(self, x):
        _ops.scale_(self._data, float(x))
        return self

    def __itruediv__(self, x):
        _ops.scale_(self._data, 1.0 / float(x))
        return self

    def normalize_(self):
        """In-place division by the norm without a host round trip."""
        _ops.scale_rsqrt_dev_(self._data, self._norm2_device())
        return self

    def __matmul__(self, other):
        """Contract all column legs of self with all row legs of other: one grouped
        GEMM over every irrep present on both sides (symmetric_tensor.py:453-489)."""
        assert type(self) is type(other)
        assert self._shape[self._nrr :] == other._shape[: other._nrr]
        assert not _debug_config["debug_checks"] or (self._signature[self._nrr :] ^ other._signature[: other._nrr]).all()
        key = (
            type(self), self._fortran, other._fortran,
            self._block_irreps.tobytes(), self._bshapes.tobytes(), self._boffs.tobytes(),
            other._block_irreps.tobytes(), other._bshapes.tobytes(), other._boffs.tobytes(),
        )
        with _MATMUL_LAYOUTS_LOCK:
            layout = _MATMUL_LAYOUTS.get(key)
            if layout is not None:
                _MATMUL_LAYOUTS.move_to_end(key)
        if layout is None:
            layout = self._matmul_layout(other)
            with _MATMUL_LAYOUTS_LOCK:
                _MATMUL_LAYOUTS[key] = layout
                if len(_MATMUL_LAYOUTS) > _MATMUL_LAYOUTS_MAX:
                    _MATMUL_LAYOUTS.popitem(last=False)
        groups, shapes, offs, total, irreps = layout
        out = _ops.empty(total)
        _ops.grouped_gemm(groups, self._data, other._data, out, split=_precision_split())
        signature = np.empty((self._nrr + other._ndim - other._nrr,), dtype=bool)
        signature[: self._nrr] = self._signature[: self._nrr]
        signature[self._nrr :] = other._signature[other._nrr :]
        shape = self._shape[: self._nrr] + other._shape[other._nrr :]
        return type(self)._build(self._row_reps, other._col_reps, out, shapes, offs, irreps, signature, shape=shape)

    def _matmul_layout(self, other):
        """Blocks present on both sides (block irreps are sorted), in irrep order:
        the grouped-GEMM groups, output block shapes, offsets, size and irreps.
        Depends only on the two block structures, so __matmul__ caches it."""
        ia, ib = self._block_irreps, other._block_irreps
        pos = np.minimum(np.searchsorted(ia, ib), max(ia.size - 1, 0))
        match = (ia[pos] == ib) if ia.size else np.zeros(ib.shape, dtype=bool)
        i_idx, j_idx = pos[match], np.flatnonzero(match)
        m, kk = self._bshapes[i_idx, 0], self._bshapes[i_idx, 1]
        n = other._bshapes[j_idx, 1]
        shapes = np.stack([m, n], axis=1)
        irreps = ia[i_idx]
        offs, total = _packed_offsets(shapes)
        kb = other._bshapes[j_idx, 0]
        assert not _debug_config["debug_checks"] or (kb == kk).all()
        one = np.ones_like(m)
        ars, acs = (one, m) if self._fortran else (kk, one)
        brs, bcs = (one, kb) if other._fortran else (n, one)
        groups = np.stack([m, n, kk, self._boffs[i_idx], ars, acs, other._boffs[j_idx], brs, bcs, offs, n], axis=1)
        return groups, shapes, offs, total, irreps

    # misc ------------------------------------------------------------------
    def match_representations(self, other):
        if (
            type(self) is not type(other)
            or self._shape != other.shape
            or (self._signature != other.signature).any()
        ):
            return False
        for r, r2 in zip(self._row_reps, other.row_reps, strict=True):
            if np.shape(r) != np.shape(r2) or (np.asarray(r) != np.asarray(r2)).any():
                return False
        for r, r2 in zip(self._col_reps, other.col_reps, strict=True):
            if np.shape(r) != np.shape(r2) or (np.asarray(r) != np.asarray(r2)).any():
                return False
        return True

    def get_row_representation(self):
        return self.combine_representations(self._row_reps, self._signature[: self._nrr])

    def get_column_representation(self):
        return self.combine_representations(self._col_reps, ~self._signature[self._nrr :])

    # transpose and permute -------------------------------------------------
    def permute(self, row_axes, col_axes):
        """Permute legs and regroup them into rows and columns (symmetric_tensor.py:166-212)."""
        row_axes = tuple(int(a) for a in row_axes)
        col_axes = tuple(int(a) for a in col_axes)
        axes = row_axes + col_axes
        if sorted(axes) != list(range(self._ndim)):
            raise ValueError("Axes do not match tensor")
        if row_axes == tuple(range(self._nrr)) and col_axes == tuple(range(self._nrr, self._ndim)):
            return self
        if row_axes == tuple(range(self._nrr, self._ndim)) and col_axes == tuple(range(self._nrr)):
            return self.transpose()
        return self._permuted(axes, len(row_axes))

    def _permuted(self, axes, nrr):
        """New tensor with legs `axes` (first nrr are rows); data moved by the
        symmetry's _permute_data."""
        signature = np.empty((self._ndim,), dtype=bool)
        reps = [None] * self._ndim
        for i, ax in enumerate(axes):
            signature[i] = self._signature[ax]
            reps[i] = self._row_reps[ax] if ax < self._nrr else self._col_reps[ax - self._nrr]
        if self._nblocks == 0:
            data, shapes, block_irreps = _ops.empty(0), np.zeros((0, 2), np.int64), self._block_irreps
        else:
            data, shapes, block_irreps = self._permute_data(axes, nrr)
        offs, _ = _packed_offsets(shapes)
        shape = tuple(self._shape[a] for a in axes)
        tp = type(self)._build(reps[:nrr], reps[nrr:], data, shapes, offs, block_irreps, signature, shape=shape)
        assert check_norm(self, tp)
        return tp

    def transpose(self):
        """Matrix transpose (zero-copy relabelling of blocks)."""
        order, block_irreps = self._transpose_order()
        s = self._signature[np.arange(-self._ndim + self._nrr, self._nrr) % self._ndim]
        return type(self)._build(
            self._col_reps,
            self._row_reps,
            self._data,
            self._bshapes[order][:, ::-1].copy(),
            self._boffs[order],
            block_irreps,
            s,
            not self._fortran,
            self._shape[self._nrr :] + self._shape[: self._nrr],
        )

    def dual(self):
        return self.dagger().conjugate().transpose()

    def conjugate(self):
        # real coefficients: conjugation is the identity (data shared like numpy views)
        return self._with()

    def dagger(self):
        """Adjoint: swap rows and columns, reverse all signatures; blocks are shared."""
        s = np.empty((self._ndim,), dtype=bool)
        s[: self._ndim - self._nrr] = ~self._signature[self._nrr :]
        s[self._ndim - self._nrr :] = ~self._signature[: self._nrr]
        return type(self)._build(
            self._col_reps,
            self._row_reps,
            self._data,
            self._bshapes[:, ::-1].copy(),
            self._boffs,
            self._block_irreps,
            s,
            not self._fortran,
            self._shape[self._nrr :] + self._shape[: self._nrr],
        )

    def contiguous(self):
        """Copy with row-major blocks packed in block order (no-op if already so)."""
        offs, total = _packed_offsets(self._bshapes)
        if not self._fortran and np.array_equal(offs, self._boffs):
            return self
        out = _ops.empty(total)
        descs = []
        for b in range(self._nblocks):
            m, n = self._bshapes[b]
            rs, cs = self._strides(b)
            descs.append((m, n, self._boffs[b], rs, cs, offs[b], n, 1, 0, 0, 0, 1.0, 0.0))
        _ops.copy2d(descs, self._data, out)
        return type(self)._build(
            self._row_reps, self._col_reps, out, self._bshapes, offs, self._block_irreps,
            self._signature, shape=self._shape,
        )

    # linear algebra ----------------------------------------------------------
    def is_square_matrix(self):
        if self._nrr != self.ndim // 2:
            return False
        if (self._signature[: self._nrr] == self._signature[self._nrr :]).any():
            return False
        for r1, r2 in zip(self.row_reps, self.col_reps, strict=True):
            if np.shape(r1) != np.shape(r2) or (np.asarray(r1) != np.asarray(r2)).any():
                return False
        return True

    def _norm2_device(self):
        if self._nblocks == 0:
            return _ops.zeros(1)
        dims = self._dims()
        sizes = self._bshapes[:, 0] * self._bshapes[:, 1]
        if np.all(dims == 1.0):
            return _ops.weighted_norm2(self._data, np.array([0, sizes.sum()]), np.ones((1,)))
        order = np.argsort(self._boffs, kind="stable")
        seg = np.empty((self._nblocks + 1,), dtype=np.int64)
        seg[:-1] = self._boffs[order]
        seg[-1] = seg[-2] + sizes[order[-1]]
        return _ops.weighted_norm2(self._data, seg, dims[order])

    def norm(self):
        """Frobenius norm sqrt(sum_b dim(irrep_b) ||b||^2) (symmetric_tensor.py:624-631)."""
        return float(np.sqrt(float(_ops.download(self._norm2_device())[0])))

    def trace(self):
        assert self.is_square_matrix()
        if self._nblocks == 0:
            return 0.0
        offs, ns, lds = [], [], []
        for b in range(self._nblocks):
            m, n = self._bshapes[b]
            assert m == n
            offs.append(self._boffs[b])
            ns.append(m)
            lds.append(n)  # diagonal stride n+1 in either memory order
        return float(_ops.download(_ops.weighted_trace(self._data, offs, ns, lds, self._dims()))[0])

    def full_contract(self, other):
        """Tr(self @ other) summed over irreps with their dimension (:639-653)."""
        assert self.match_representations(other.dagger())
        prod = self @ other
        return prod.trace() if prod.nblocks else 0.0

    def _diagonal_mul(self, dt, *, left=False):
        """Scale columns (right) or rows (left) by DiagonalTensor weights
        (symmetric_tensor.py:1189-1248)."""
        diag_irreps = np.asarray(dt.block_irreps)
        vals, voffs = dt._device_values()
        idx = np.arange(len(diag_irreps))
        if (left and self._signature[0]) or (not left and not self._signature[-1]):
            conj = np.array([self.conjugate_irrep(r) for r in diag_irreps])
            idx = conj.argsort()
            diag_irreps = conj[idx]
        if left:
            assert self._nrr == 1
        else:
            assert self._ndim - self._nrr == 1
        dpos = {int(r): int(i) for r, i in zip(diag_irreps, idx, strict=True)}
        keep = []
        for b, r in enumerate(self._block_irreps):
            if int(r) in dpos:
                keep.append((b, dpos[int(r)]))
            else:
                print("Warning: missing block in diagonal blocks")
        shapes = self._bshapes[[b for b, _ in keep]].reshape(-1, 2)
        offs, total = _packed_offsets(shapes)
        out = _ops.empty(total)
        descs = []
        for k, (b, di) in enumerate(keep):
            m, n = self._bshapes[b]
            rs, cs = self._strides(b)
            w_off = voffs[di]
            descs.append(
                (m, n, self._boffs[b], rs, cs, offs[k], n, 1, w_off if left else 0,
                 0 if left else w_off, 0, 1.0, 0.0)
            )
        if left:
            _ops.copy2d(descs, self._data, out, row_w=vals)
        else:
            _ops.copy2d(descs, self._data, out, col_w=vals)
        irreps = self._block_irreps[[b for b, _ in keep]]
        return type(self)._build(
            self._row_reps, self._col_reps, out, shapes, offs, irreps, self._signature, shape=self._shape
        )

    # --- batched SVD core ---------------------------------------------------
    def _svd_blocks(self, block_ids):
        """Full thin SVD of the selected blocks in one batched launch.
        Returns (U buffer, u offsets, S buffer, s offsets, VT buffer, vt offsets, q list)."""
        uo, so, vo, wo, qs, descs = [], [], [], [], [], []
        nu = ns = nv = nw = 0
        for b in block_ids:
            m, n = (int(x) for x in self._bshapes[b])
            q = min(m, n)
            if q > _ops.SVD_MAX_MIN_DIM:
                msg = f"block of shape {(m, n)} exceeds the dense device SVD limit"
                raise NotImplementedError(msg)
            rs, cs = self._strides(b)
            uo.append(nu)
            so.append(ns)
            vo.append(nv)
            wo.append(nw)
            qs.append(q)
            descs.append((m, n, self._boffs[b], rs, cs, nu, ns, nv, nw))
            nu += m * q
            ns += q
            nv += q * n
            nw += q * (max(m, n) + q)
        U = _ops.empty(nu)
        S = _ops.empty(ns)
        VT = _ops.empty(nv)
        work = _ops.empty(nw)
        _ops.batched_svd(descs, self._data, U, S, VT, work)
        return U, uo, S, so, VT, vo, qs

    def svd(self):
        """Blockwise thin SVD: U (rows | mid), s, V (mid | cols) (symmetric_tensor.py:691-710)."""
        bids = list(range(self._nblocks))
        U, uo, S, so, VT, vo, qs = self._svd_blocks(bids)
        s_host = _ops.download(S)
        s_blocks = [s_host[so[i] : so[i] + qs[i]].copy() for i in range(len(bids))]
        degen = np.array(qs, dtype=np.int64)
        mid_rep = self.init_representation(degen, self._block_irreps)
        usign = np.ones((self._nrr + 1,), dtype=bool)
        usign[: self._nrr] = self._signature[: self._nrr]
        ushapes = np.array([(self._bshapes[b, 0], qs[i]) for i, b in enumerate(bids)], np.int64).reshape(-1, 2)
        Ut = type(self)._build(self._row_reps, (mid_rep,), U, ushapes, np.array(uo, np.int64), self._block_irreps, usign)
        degens = [self.irrep_dimension(irr) for irr in self._block_irreps]
        s = DiagonalTensor(s_blocks, mid_rep, self._block_irreps, degens, self._symmetry)
        vsign = np.zeros((self._ndim - self._nrr + 1,), dtype=bool)
        vsign[1:] = self._signature[self._nrr :]
        vshapes = np.array([(qs[i], self._bshapes[b, 1]) for i, b in enumerate(bids)], np.int64).reshape(-1, 2)
        Vt = type(self)._build((mid_rep,), self._col_reps, VT, vshapes, np.array(vo, np.int64), self._block_irreps, vsign)
        return Ut, s, Vt

    def _truncate(self, kept, vals_blocks, block_irreps, *, left=None, right=None):
        """Build (U, s, V) from kept counts. left/right = (buffer, offsets, rows|cols,
        q, col_idx per block or None); only the first kept[b] vectors are kept."""
        nonempty = [b for b in range(len(kept)) if kept[b] > 0]
        nkept = np.array([kept[b] for b in nonempty], dtype=np.int64)
        kept_irreps = np.array([block_irreps[b] for b in nonempty], dtype=block_irreps.dtype)
        s_blocks = [np.asarray(vals_blocks[b][: kept[b]]) for b in nonempty]
        s_rep = self.init_representation(nkept, kept_irreps)
        degens = [self.irrep_dimension(irr) for irr in kept_irreps]
        s = DiagonalTensor(s_blocks, s_rep, kept_irreps, degens, self._symmetry)
        u = v = None
        if left is not None:
            buf, offs, rows, qs, cidx = left
            shapes = np.array([(rows[b], kept[b]) for b in nonempty], np.int64).reshape(-1, 2)
            doffs, total = _packed_offsets(shapes)
            out = _ops.empty(total)
            descs = []
            idx_dev = None
            if cidx is not None:
                idx_host = np.concatenate([cidx[b][: kept[b]] for b in nonempty]).astype(np.int64) if nonempty else np.zeros((0,), np.int64)
                idx_dev = _ops.upload(idx_host, dtype=torch.int64)
            ci = 0
            for k, b in enumerate(nonempty):
                m, c = shapes[k]
                descs.append((m, c, offs[b], qs[b], 1, doffs[k], c, 1, 0, 0, ci, 1.0, 0.0))
                ci += c
            _ops.copy2d(descs, buf, out, col_idx=idx_dev)
            usign = np.hstack((self._signature[: self._nrr], [True]))
            u = type(self)._build(self._row_reps, (s_rep,), out, shapes, doffs, kept_irreps, usign)
        if right is not None:
            buf, offs, cols, qs, _ = right
            shapes = np.array([(kept[b], cols[b]) for b in nonempty], np.int64).reshape(-1, 2)
            doffs, total = _packed_offsets(shapes)
            out = _ops.empty(total)
            descs = []
            for k, b in enumerate(nonempty):
                r, n = shapes[k]
                descs.append((r, n, offs[b], n, 1, doffs[k], n, 1, 0, 0, 0, 1.0, 0.0))
            _ops.copy2d(descs, buf, out)
            vsign = np.hstack(([False], self._signature[self._nrr :]))
            v = type(self)._build((s_rep,), self._col_reps, out, shapes, doffs, kept_irreps, vsign)
        return u, s, v

    def truncated_svd(self, cut, *, dmax_full=100, oversampling=0, atol=0.0, rtol=None, degen_ratio=1.0):
        """Blockwise SVD then global multiplet-aware truncation to about `cut` values
        (symmetric_tensor.py:835-877). Every block is decomposed exactly on the device
        (batched one-sided Jacobi), which covers both of the reference's dense and
        ARPACK branches; dmax_full and oversampling are accepted for compatibility."""
        del dmax_full, oversampling
        bids = list(range(self._nblocks))
        U, uo, S, so, VT, vo, qs = self._svd_blocks(bids)
        s_host = _ops.download(S)
        raw_s = [s_host[so[i] : so[i] + qs[i]] for i in range(len(bids))]
        dims = np.array([self.irrep_dimension(r) for r in self._block_irreps])
        cuts = find_block_cuts(raw_s, cut, dims=dims, atol=atol, rtol=rtol, degen_ratio=degen_ratio)
        rows = [int(self._bshapes[b, 0]) for b in bids]
        cols = [int(self._bshapes[b, 1]) for b in bids]
        return self._truncate(
            cuts, raw_s, self._block_irreps,
            left=(U, uo, rows, qs, None), right=(VT, vo, cols, qs, None),
        )

    def pinv(self, *, atol=0.0, rtol=None):
        u, s, v = self.truncated_svd(2**31, atol=atol, rtol=rtol)
        return v.dagger() * (1 / s) @ u.dagger()

    def qr(self):
        """Economic QR of every block in one batched Householder launch, LAPACK
        conventions (symmetric_tensor.py:655-671): Q (rows | mid), R (mid | cols)."""
        from ..misc_tools import block_linalg as bl

        A = bl.Mats.from_tensor_blocks(self, range(self._nblocks))
        Q, R = bl.qr(A)
        degen = np.minimum(self._bshapes[:, 0], self._bshapes[:, 1])
        mid_rep = self.init_representation(degen, self._block_irreps)
        qsign = np.ones((self._nrr + 1,), dtype=bool)
        qsign[: self._nrr] = self._signature[: self._nrr]
        rsign = np.zeros((self._ndim - self._nrr + 1,), dtype=bool)
        rsign[1:] = self._signature[self._nrr :]
        Qt = type(self)._build(
            self._row_reps, (mid_rep,), Q.buf, np.stack([Q.rows, Q.cols], 1), Q.off,
            self._block_irreps, qsign,
        )
        Rt = type(self)._build(
            (mid_rep,), self._col_reps, R.buf, np.stack([R.rows, R.cols], 1), R.off,
            self._block_irreps, rsign,
        )
        return Qt, Rt

    def rq(self):
        """A = R Q with Q orthonormal rows, from the QR of the transpose: R is lower
        triangular here (LAPACK's rq returns an upper triangular R)."""
        q, r = self.transpose().qr()
        return r.transpose(), q.transpose()

    def _shifted_svd(self):
        """SVD of every block shifted by sigma*I with sigma = ||A||_F, which makes each
        (symmetric) block positive semi-definite so that its SVD is its eigen
        decomposition: lambda = s - sigma, eigenvectors = right singular vectors."""
        sigma = self.norm()
        if sigma == 0.0:
            sigma = 1.0
        shifted = self.contiguous()
        if shifted is self:
            shifted = self.copy()
        shift = _ops.upload(np.array([sigma]))
        descs = []
        for b in range(self._nblocks):
            m, n = self._bshapes[b]
            descs.append((m, 1, 0, 0, 0, shifted._boffs[b], n + 1, 1, 0, 0, 0, 1.0, 1.0))
        _ops.copy2d(descs, shift, shifted._data)
        U, uo, S, so, VT, vo, qs = shifted._svd_blocks(list(range(self._nblocks)))
        return sigma, S, so, VT, vo, qs

    def truncated_eigh(self, nvals, *, atol=0.0, rtol=None, degen_ratio=1.0, dmax_full=100,
                       compute_vectors=True, rng=None, maxiter=4000, arpack_tol=0, matmat=None):
        """Largest-magnitude eigenpairs of a real symmetric block matrix, multiplet-aware
        truncation to about nvals values (symmetric_tensor.py:1087-1132, 1250-1434).
        Every present block is diagonalised exactly on the device; missing blocks
        contribute zero eigenvalues that the cut removes."""
        if matmat is not None:
            return self._sparse_eig(
                matmat, nvals, symmetric=True, atol=atol, rtol=rtol, degen_ratio=degen_ratio,
                dmax_full=dmax_full, compute_vectors=compute_vectors, rng=rng, maxiter=maxiter,
                arpack_tol=arpack_tol,
            )
        if not self.is_square_matrix():
            raise ValueError("Tensor legs do not define a square matrix")
        all_irreps, all_shapes = self.get_block_sizes(self._row_reps, self._col_reps, self._signature)
        sigma, S, so, VT, vo, qs = self._shifted_svd()
        s_host = _ops.download(S) if self._nblocks else np.zeros((0,))
        present = {int(r): i for i, r in enumerate(self._block_irreps)}
        nb = len(all_irreps)
        dims = np.array([self.irrep_dimension(r) for r in all_irreps], dtype=np.int64)
        val_blocks = [None] * nb
        abs_blocks = [np.zeros((1,))] * nb
        order_blocks = [None] * nb
        vec_offs = [0] * nb
        vec_q = [0] * nb
        rows = [int(all_shapes[b, 0]) for b in range(nb)]
        for bi in range(nb):
            k = nvals // dims[bi] + 1
            i = present.get(int(all_irreps[bi]))
            if i is None:
                continue
            lam = s_host[so[i] : so[i] + qs[i]] - sigma
            absl = np.abs(lam)
            so_ = absl.argsort()[: -k - 1 : -1]
            val_blocks[bi] = lam[so_]
            abs_blocks[bi] = absl[so_]
            order_blocks[bi] = so_
            vec_offs[bi] = vo[i]
            vec_q[bi] = qs[i]
        cuts = find_block_cuts(abs_blocks, nvals, dims=dims, atol=atol, rtol=rtol, degen_ratio=degen_ratio)
        for bi in range(nb):
            if val_blocks[bi] is None:
                cuts[bi] = 0
        vals = [v if v is not None else np.zeros((0,)) for v in val_blocks]
        if not compute_vectors:
            _, s, _ = self._truncate(cuts, vals, np.asarray(all_irreps))
            return s
        # eigenvectors are the rows of VT (q x n): read them transposed (s_rs=1, s_cs=n)
        left = (VT, vec_offs, rows, vec_q, order_blocks)
        u, s, _ = self._truncate_vectors_from_rows(cuts, vals, np.asarray(all_irreps), left)
        return s, u

    def _truncate_vectors_from_rows(self, kept, vals_blocks, block_irreps, left):
        """Like _truncate(left=...) with vectors stored as rows (VT layout)."""
        buf, offs, rows, qs, order = left
        nonempty = [b for b in range(len(kept)) if kept[b] > 0]
        nkept = np.array([kept[b] for b in nonempty], dtype=np.int64)
        kept_irreps = np.array([block_irreps[b] for b in nonempty], dtype=block_irreps.dtype)
        s_blocks = [np.asarray(vals_blocks[b][: kept[b]]) for b in nonempty]
        s_rep = self.init_representation(nkept, kept_irreps)
        degens = [self.irrep_dimension(irr) for irr in kept_irreps]
        s = DiagonalTensor(s_blocks, s_rep, kept_irreps, degens, self._symmetry)
        shapes = np.array([(rows[b], kept[b]) for b in nonempty], np.int64).reshape(-1, 2)
        doffs, total = _packed_offsets(shapes)
        out = _ops.empty(total)
        idx_host = (
            np.concatenate([order[b][: kept[b]] for b in nonempty]).astype(np.int64)
            if nonempty else np.zeros((0,), np.int64)
        )
        idx_dev = _ops.upload(idx_host, dtype=torch.int64)
        descs = []
        ci = 0
        for k, b in enumerate(nonempty):
            m, c = shapes[k]
            # element (i, j) of the eigenvector matrix = VT[j, i]
            descs.append((m, c, offs[b], 1, rows[b], doffs[k], c, 1, 0, 0, ci, 1.0, 0.0))
            ci += c
        _ops.copy2d(descs, buf, out, col_idx=idx_dev)
        usign = np.hstack((self._signature[: self._nrr], [True]))
        u = type(self)._build(self._row_reps, (s_rep,), out, shapes, doffs, kept_irreps, usign)
        return u, s, None

    def eigh(self, *, compute_vectors=True):
        if not self.is_square_matrix():
            raise ValueError("Tensor legs do not define a square matrix")
        return self.truncated_eigh(2**31, compute_vectors=compute_vectors)

    # Taylor degree of expm: with the scaled matrix's norm <= 1/2, the series
    # remainder is below e^(1/2) 0.5^17 / 17! ~ 2e-20 of the norm.
    _EXPM_DEGREE = 16

    def expm(self):
        """exp of every block (symmetric_tensor.py:818-824 calls scipy's expm
        on each block; missing blocks stay missing). Scaling and squaring of the
        truncated Taylor series (Moler & Van Loan, "Nineteen dubious ways",
        method 3): A is scaled by 2^-s until its norm is at most 1/2, the
        degree-16 series is evaluated by Horner's rule, T = I + A T / k for
        k = 16 .. 1, and squared s times. Every product is one grouped GEMM on
        the device, so general (non-symmetric) blocks need no host round trip.
        The tensor norm bounds every block's spectral norm."""
        if not self.is_square_matrix():
            raise ValueError("Tensor legs do not define a square matrix")
        if self.nblocks == 0:
            return self.copy()
        eye = type(self)(
            self._row_reps, self._col_reps, [np.eye(int(m)) for m, _ in self._bshapes],
            self._block_irreps, self._signature,
        )
        nrm = self.norm()
        s = max(0, int(np.ceil(np.log2(nrm / 0.5)))) if nrm > 0 else 0
        a = self._affine(2.0**-s)
        t = eye
        for k in range(self._EXPM_DEGREE, 0, -1):
            t = eye + (a @ t)._affine(1.0 / k)
        for _ in range(s):
            t = t @ t
        return t

    def solve(self, b):
        raise NotImplementedError("solve is not provided by the device engine")

    def eig(self, *, compute_vectors=True):
        """All eigenvalues of every block (symmetric_tensor.py:779-815); exact
        zeros of missing blocks are truncated as in the reference."""
        if not self.is_square_matrix():
            raise ValueError("Tensor legs do not define a square matrix")
        return self._sparse_eig(None, 2**31, symmetric=False, compute_vectors=compute_vectors)

    @classmethod
    def eigs(cls, matmat, nvals, reps, signature, *, atol=0.0, rtol=None, degen_ratio=1.0, dmax_full=100,
             compute_vectors=True, rng=None, maxiter=4000, arpack_tol=0):
        """nvals largest-magnitude eigenvalues of the real square matrix defined
        by matmat on tensors with row representations reps and signature
        (symmetric_tensor.py:880-956)."""
        m = cls.init_square_matrix(reps, signature)
        return m.truncated_eig(
            nvals, atol=atol, rtol=rtol, degen_ratio=degen_ratio, dmax_full=dmax_full,
            compute_vectors=compute_vectors, rng=rng, maxiter=maxiter, arpack_tol=arpack_tol, matmat=matmat,
        )

    @classmethod
    def eigsh(cls, matmat, nvals, reps, signature, *, atol=0.0, rtol=None, degen_ratio=1.0, dmax_full=100,
              compute_vectors=True, rng=None, maxiter=4000, arpack_tol=0):
        """Real symmetric counterpart of eigs (symmetric_tensor.py:1006-1085)."""
        m = cls.init_square_matrix(reps, signature)
        return m.truncated_eigh(
            nvals, atol=atol, rtol=rtol, degen_ratio=degen_ratio, dmax_full=dmax_full,
            compute_vectors=compute_vectors, rng=rng, maxiter=maxiter, arpack_tol=arpack_tol, matmat=matmat,
        )

    def truncated_eig(self, nvals, *, atol=0.0, rtol=None, degen_ratio=1.0, dmax_full=100,
                      compute_vectors=True, rng=None, maxiter=4000, arpack_tol=0, matmat=None):
        """nvals largest-magnitude eigenvalues of a real square block matrix, or of
        the matrix implicitly defined by matmat (symmetric_tensor.py:958-1004)."""
        return self._sparse_eig(
            matmat, nvals, symmetric=False, atol=atol, rtol=rtol, degen_ratio=degen_ratio,
            dmax_full=dmax_full, compute_vectors=compute_vectors, rng=rng, maxiter=maxiter,
            arpack_tol=arpack_tol,
        )

    def _sparse_eig(self, matmat, nvals, *, symmetric, atol=0.0, rtol=None, degen_ratio=1.0, dmax_full=100,
                    compute_vectors=False, rng=None, maxiter=4000, arpack_tol=0):
        """Blockwise largest-magnitude eigenpairs (symmetric_tensor.py:1250-1434).

        Per block of the square matrix, k = nvals // irrep_dim + 1 values are
        wanted. Blocks smaller than max(dmax_full, 3k) are formed explicitly (from
        self, or with one matmat on identity blocks, on the device) and
        diagonalized in full with LAPACK on the host: they are at most a few
        hundred wide. Larger blocks go to the device Krylov-Schur solver, which
        applies matmat to one-block tensors whose data are the Krylov vectors
        themselves. Then the multiplet-aware global cut of find_block_cuts."""
        from ..misc_tools import krylov

        if rng is None:
            rng = np.random.default_rng()
        if not self.is_square_matrix():
            raise ValueError("Tensor legs do not define a square matrix")
        use_blocks = matmat is None
        irreps, shapes = self.get_block_sizes(self._row_reps, self._col_reps, self._signature)
        irreps = np.asarray(irreps)
        nb = len(irreps)
        dims = np.array([self.irrep_dimension(r) for r in irreps], dtype=np.int64)
        sigv = np.hstack((self._signature[: self._nrr], [True]))
        vals = [np.zeros((0,))] * nb
        absv = [np.zeros((1,))] * nb
        vecs = [None] * nb  # host (d x k) array or device (buffer, d, k) row-major
        ks = [int(nvals // dims[b] + 1) for b in range(nb)]
        full = [b for b in range(nb) if shapes[b, 0] < max(dmax_full, 3 * ks[b])]
        sparse = [b for b in range(nb) if b not in set(full)]

        def keep(bi, lam, vec):
            order = np.argsort(-np.abs(lam), kind="stable")[: ks[bi]]
            vals[bi] = lam[order]
            absv[bi] = np.abs(lam[order])
            if vec is not None:
                vecs[bi] = vec[:, order]

        # dense blocks
        if full:
            if use_blocks:
                st = self
            else:
                irr_full = np.ascontiguousarray(irreps[full])
                rfull = self.init_representation(shapes[full, 0], irr_full)
                eyes = [np.eye(int(shapes[b, 0])) for b in full]
                st = matmat(type(self)(self._row_reps, (rfull,), eyes, irr_full, sigv))
            present = {int(r): j for j, r in enumerate(st.block_irreps)}
            host_blocks = st.blocks  # one readback for every dense block
            for bi in full:
                j = present.get(int(irreps[bi]))
                if j is None:
                    continue  # missing block: zero eigenvalues, truncated
                b = np.ascontiguousarray(host_blocks[j])
                if symmetric:
                    lam, vec = np.linalg.eigh(b) if compute_vectors else (np.linalg.eigvalsh(b), None)
                else:
                    lam, vec = np.linalg.eig(b) if compute_vectors else (np.linalg.eigvals(b), None)
                keep(bi, lam, vec)

        # large blocks: device Krylov-Schur
        for bi in sparse:
            d = int(shapes[bi, 0])
            v0 = _ops.upload(rng.standard_normal(size=(d,)))
            irr1 = np.ascontiguousarray(irreps[bi : bi + 1])
            if use_blocks:
                js = [j for j, r in enumerate(self._block_irreps) if int(r) == int(irreps[bi])]
                if not js:
                    continue
                j = js[0]
                rs, cs = self._strides(j)
                off = int(self._boffs[j])

                def apply(x, y, off=off, rs=rs, cs=cs, d=d):
                    _ops.grouped_gemm([(d, 1, d, off, rs, cs, 0, 1, 1, 0, 1)], self._data, x, y)
            else:
                brep = self.init_representation(np.ones((1,), dtype=np.int64), irr1)

                def probe(x, brep=brep, irr1=irr1, d=d):
                    st0 = type(self)._build(self._row_reps, (brep,), x, np.array([[d, 1]]), np.zeros(1, np.int64),
                                            irr1, sigv)
                    st1 = matmat(st0)
                    js = [j for j, r in enumerate(st1.block_irreps) if int(r) == int(irr1[0])]
                    return st1, (js[0] if js else None)

                _, j0 = probe(v0)
                if j0 is None:
                    continue  # the block does not appear in the output: zero block

                def apply(x, y, probe=probe, d=d):
                    st1, j = probe(x)
                    if j is None or tuple(st1._bshapes[j]) != (d, 1):
                        raise ValueError("matmat changed the block structure of its input")
                    _ops.copy2d([(1, d, int(st1._boffs[j]), 1, 1, 0, 1, 1, 0, 0, 0, 1.0, 0.0)], st1._data, y)
            k = ks[bi]
            try:
                lam, X = krylov.largest_eigenpairs(
                    apply, d, k, v0, maxiter=maxiter, tol=arpack_tol, compute_vectors=compute_vectors,
                    symmetric=symmetric,
                )
            except krylov.KrylovNoConvergence as err:
                print("Warning: Krylov-Schur did not converge", err)
                lam, X = err.eigenvalues, err.eigenvectors
                print(f"Keep {lam.size}/{k} converged values")
            vals[bi] = lam
            absv[bi] = np.abs(lam) if lam.size else np.zeros((1,))
            vecs[bi] = None if X is None else (X, d, lam.size)

        cuts = find_block_cuts(absv, nvals, dims=dims, atol=atol, rtol=rtol, degen_ratio=degen_ratio)
        for bi in range(nb):
            cuts[bi] = min(int(cuts[bi]), vals[bi].size)
        if symmetric:
            vals = [np.real(v) for v in vals]
        if not compute_vectors:
            _, s, _ = self._truncate(cuts, vals, irreps)
            return s
        return self._eig_vectors(cuts, vals, irreps, vecs)

    def _eig_vectors(self, cuts, vals, irreps, vecs):
        """Pack the kept eigenvectors of every block into one device buffer and
        build (s, u) with _truncate."""
        rows, qs, offs = [], [], []
        total = 0
        host_parts = []
        dev_parts = []
        for bi, v in enumerate(vecs):
            c = int(cuts[bi])
            if c == 0 or v is None:
                rows.append(0)
                qs.append(0)
                offs.append(total)
                continue
            if isinstance(v, tuple):
                X, d, q = v
                dev_parts.append((X, d, q, total))
            else:
                if np.iscomplexobj(v) and np.abs(v[:, :c].imag).max(initial=0.0) > 0:
                    raise NotImplementedError("complex eigenvectors are not supported by the float64 device tensors")
                d, q = v.shape
                host_parts.append((np.ascontiguousarray(np.real(v)), total))
            rows.append(d)
            qs.append(q)
            offs.append(total)
            total += d * q
        if any(np.iscomplexobj(v[: int(c)]) and np.abs(np.imag(v[: int(c)])).max(initial=0.0) > 0
               for v, c in zip(vals, cuts, strict=True)):
            raise NotImplementedError("complex eigenvectors are not supported by the float64 device tensors")
        host = np.zeros((total,))
        for arr, o in host_parts:
            host[o : o + arr.size] = arr.reshape(-1)
        buf = _ops.upload(host)
        for X, d, q, o in dev_parts:
            _ops.copy2d([(1, d * q, 0, 1, 1, o, 1, 1, 0, 0, 0, 1.0, 0.0)], X, buf)
        vals = [np.real(v) for v in vals]
        u, s, _ = self._truncate(cuts, vals, irreps, left=(buf, offs, rows, qs, None))
        return s, u

    @classmethod
    def init_square_matrix(cls, reps, signature):
        s = np.asarray(signature, dtype=bool)
        return cls(reps, reps, [], np.zeros((0,), dtype=cls.singlet().dtype), np.hstack((s, ~s)))

    # I/O (npz schema of symmetric_tensor.py:1137-1184) ----------------------
    def save_to_file(self, savefile):
        np.savez_compressed(savefile, **self.get_data_dic())

    def get_data_dic(self, *, prefix=""):
        data = {
            prefix + "_symmetry": self._symmetry,
            prefix + "_n_row_reps": self._nrr,
            prefix + "_n_col_reps": self._ndim - self._nrr,
            prefix + "_block_irreps": self._block_irreps,
            prefix + "_signature": self._signature,
        }
        for ri, r in enumerate(self._row_reps):
            data[f"{prefix}_row_rep_{ri}"] = r
        for ci, c in enumerate(self._col_reps):
            data[f"{prefix}_col_rep_{ci}"] = c
        for bi, b in enumerate(self.blocks):
            data[f"{prefix}_block_{bi}"] = np.ascontiguousarray(b)
        return data

    @classmethod
    def load_from_dic(cls, data, *, prefix=""):
        if cls._symmetry != data[prefix + "_symmetry"][()]:
            raise ValueError(f"Saved SymmetricTensor does not match type {cls}")
        nrr = int(data[prefix + "_n_row_reps"])
        row_reps = [data[f"{prefix}_row_rep_{ri}"] for ri in range(nrr)]
        nrc = int(data[prefix + "_n_col_reps"])
        col_reps = [data[f"{prefix}_col_rep_{ri}"] for ri in range(nrc)]
        block_irreps = data[prefix + "_block_irreps"]
        signature = data[prefix + "_signature"]
        blocks = [data[f"{prefix}_block_{bi}"] for bi in range(block_irreps.size)]
        return cls(row_reps, col_reps, blocks, block_irreps, signature)

    @classmethod
    def load_from_file(cls, savefile, *, prefix=""):
        with np.load(savefile) as fin:
            return cls.load_from_dic(fin, prefix=prefix)


# Precision form of the block GEMMs ("fp64" or "split"); process-wide switch.
_PRECISION = {"gemm": "fp64"}

# grouped-GEMM layouts of __matmul__ by the operands' block structures (LRU;
# serving slots call __matmul__ from several host threads)
_MATMUL_LAYOUTS = OrderedDict()
_MATMUL_LAYOUTS_MAX = 4096
_MATMUL_LAYOUTS_LOCK = threading.Lock()


def set_gemm_precision(mode):
    """'fp64' (default, FP64 MFMA) or 'split' (FP32 MFMA hi/lo compensated products)."""
    if mode not in ("fp64", "split"):
        raise ValueError("mode must be 'fp64' or 'split'")
    _PRECISION["gemm"] = mode


def get_gemm_precision():
    return _PRECISION["gemm"]


def _precision_split():
    return _PRECISION["gemm"] == "split"
