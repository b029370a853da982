"""Oracle: abelian (U(1) / Z2-as-int8 / trivial) block-sparse tensor in numpy.

Blocks are numpy arrays, one per irrep of the fused row charges, sorted by
irrep. Follows frostspin/symmetric_tensor/{symmetric_tensor,abelian_symmetric_tensor,
u1_symmetric_tensor}.py.

Pinned by (tests/test_oracle.py):
* fuse: int8 closed form on small cases (Z2 {-128,0} wraps), brute-force
  Kronecker sum;
* from_array/toarray: round trip equality and the reference's stable in-block
  row order (brute force on tiny reps);
* permute: exact equality with numpy dense transpose (test_u1_tensor.py:59-77);
* matmul: dense matrix product;
* find_block_cuts: hand-computed cases (degenerate multiplets, cutoff,
  dims weighting) and the reference's docstring rules;
* truncated_svd/eigh: singular values / eigenvalues of the dense matrix
  (scipy), truncation error = norm of discarded values.
"""

import numpy as np
import scipy.linalg as lg


def fuse(reps, signature):
    """Kronecker sum of (+/-)charges, row-major, int8 wrap-around
    (u1_symmetric_tensor.py:8-19, 35-40)."""
    out = np.zeros((1,), dtype=np.int16)
    for r, s in zip(reps, signature, strict=True):
        c = np.asarray(r, dtype=np.int16)
        if s:
            c = -c
        out = (out[:, None] + c[None, :]).reshape(-1)
    return ((out + 128) % 256 - 128).astype(np.int8)


def find_block_cuts(block_values, cut_target, dims=None, atol=0.0, rtol=None, degen_ratio=1.0):
    """Greedy multiplet-aware truncation (misc_tools/svd_tools.py:7-74)."""
    vals = [np.asarray(v, dtype=np.float64) for v in block_values]
    n = len(vals)
    dims = np.ones((n,), dtype=int) if dims is None else np.asarray(dims, dtype=int)
    if rtol is None:
        rtol = sum(len(v) for v in vals) * np.finfo(np.float64).eps
    cutoff = atol + rtol * max(v[0] for v in vals)
    heads = np.array([v[0] if v.size else -1.0 for v in vals])
    cuts = np.zeros((n,), dtype=np.int64)
    kept = 0
    bi = int(heads.argmax())
    while heads[bi] > cutoff and kept < cut_target:
        c = degen_ratio * heads[bi]
        while heads[bi] >= c:
            cuts[bi] += 1
            kept += dims[bi]
            heads[bi] = vals[bi][cuts[bi]] if cuts[bi] < vals[bi].size else -1.0
            bi = int(heads.argmax())
    return cuts


class Tensor:
    """Abelian tensor: reps are int8 charge arrays (trivial symmetry: all zero)."""

    def __init__(self, row_reps, col_reps, blocks, block_irreps, signature):
        self.row_reps = tuple(np.asarray(r, dtype=np.int8) for r in row_reps)
        self.col_reps = tuple(np.asarray(r, dtype=np.int8) for r in col_reps)
        self.blocks = tuple(np.asarray(b, dtype=np.float64) for b in blocks)
        self.block_irreps = np.asarray(block_irreps, dtype=np.int8)
        self.signature = np.asarray(signature, dtype=bool)
        self.nrr = len(self.row_reps)
        self.shape = tuple(r.size for r in self.row_reps + self.col_reps)
        self.ndim = len(self.shape)

    @property
    def nblocks(self):
        return len(self.blocks)

    def row_charges(self):
        return fuse(self.row_reps, self.signature[: self.nrr])

    def col_charges(self):
        return fuse(self.col_reps, ~self.signature[self.nrr :])

    @staticmethod
    def block_sizes(row_reps, col_reps, signature):
        """(abelian_symmetric_tensor.py:287-318)"""
        rc = fuse(row_reps, signature[: len(row_reps)])
        cc = fuse(col_reps, ~np.asarray(signature)[len(row_reps) :])
        ri, rn = np.unique(rc, return_counts=True)
        ci, cn = np.unique(cc, return_counts=True)
        a, b = (ri[:, None] == ci).nonzero()
        return ri[a], np.array([rn[a], cn[b]]).T.reshape(-1, 2)

    @classmethod
    def random(cls, row_reps, col_reps, signature, rng):
        """Per-block rng.random(shape) - 0.5 in block order (symmetric_tensor.py:514-546)."""
        irr, shapes = cls.block_sizes(row_reps, col_reps, signature)
        blocks = [rng.random(tuple(sh)) - 0.5 for sh in shapes]
        return cls(row_reps, col_reps, blocks, irr, signature)

    @classmethod
    def from_array(cls, arr, row_reps, col_reps, signature):
        """Blocks = arr[rows of irrep][:, cols of irrep], rows in stable order
        (abelian_symmetric_tensor.py:19-40, 363-373)."""
        nrr = len(row_reps)
        signature = np.asarray(signature, dtype=bool)
        rc = fuse(row_reps, signature[:nrr])
        cc = fuse(col_reps, ~signature[nrr:])
        m = np.asarray(arr, dtype=np.float64).reshape(rc.size, cc.size)
        blocks, irreps = [], []
        for irr in np.unique(rc):
            ci = (cc == irr).nonzero()[0]
            if ci.size:
                ri = (rc == irr).nonzero()[0]
                blocks.append(m[ri][:, ci])
                irreps.append(irr)
        return cls(row_reps, col_reps, blocks, irreps, signature)

    def toarray(self, as_matrix=False):
        """(abelian_symmetric_tensor.py:63-75)"""
        rc, cc = self.row_charges(), self.col_charges()
        m = np.zeros((rc.size, cc.size))
        for irr, b in zip(self.block_irreps, self.blocks, strict=True):
            m[np.ix_((rc == irr).nonzero()[0], (cc == irr).nonzero()[0])] = b
        return m if as_matrix else m.reshape(self.shape)

    def permute(self, row_axes, col_axes):
        """Scatter each old coefficient to divmod(ori[i] + oci[j], ncols)
        (abelian_symmetric_tensor.py:103-250, 395-430)."""
        axes = tuple(row_axes) + tuple(col_axes)
        nrr = len(row_axes)
        reps = [(self.row_reps + self.col_reps)[a] for a in axes]
        sig = self.signature[list(axes)]
        new = Tensor(reps[:nrr], reps[nrr:], [], [], sig)
        nrc, ncc = new.row_charges(), new.col_charges()
        shape = np.array(self.shape, dtype=np.int64)
        # strides of the new dense flat index for every old leg
        new_strides = np.ones((self.ndim,), dtype=np.int64)
        for i in range(self.ndim - 1, 0, -1):
            new_strides[axes[i - 1]] = new_strides[axes[i]] * shape[axes[i]]
        rsh = shape[: self.nrr]
        csh = shape[self.nrr :]
        rstr1 = np.array([np.prod(rsh[i + 1 :]) for i in range(self.nrr)], dtype=np.int64)
        cstr1 = np.array([np.prod(csh[i + 1 :]) for i in range(self.ndim - self.nrr)], dtype=np.int64)
        uniq = np.unique(nrc)
        rpos = np.zeros(nrc.size, dtype=np.int64)
        rblk = np.zeros(nrc.size, dtype=np.int64)
        for k, irr in enumerate(uniq):
            idx = (nrc == irr).nonzero()[0]
            rpos[idx] = np.arange(idx.size)
            rblk[idx] = k
        cpos = np.zeros(ncc.size, dtype=np.int64)
        ccount = np.zeros(uniq.size, dtype=np.int64)
        for k, irr in enumerate(uniq):
            idx = (ncc == irr).nonzero()[0]
            cpos[idx] = np.arange(idx.size)
            ccount[k] = idx.size
        rcount = np.array([(nrc == irr).sum() for irr in uniq])
        new_blocks = [np.zeros((rcount[k], ccount[k])) for k in range(uniq.size)]
        orc, occ = self.row_charges(), self.col_charges()
        for irr, b in zip(self.block_irreps, self.blocks, strict=True):
            ori = (orc == irr).nonzero()[0]
            oci = (occ == irr).nonzero()[0]
            ori = (ori[:, None] // rstr1 % rsh * new_strides[: self.nrr]).sum(axis=1)
            oci = (oci[:, None] // cstr1 % csh * new_strides[self.nrr :]).sum(axis=1)
            nr, nc = np.divmod(ori[:, None] + oci[None, :], ncc.size)
            kb = rblk[nr]
            for k in np.unique(kb):
                sel = kb == k
                new_blocks[k][rpos[nr[sel]], cpos[nc[sel]]] = b[sel]
        keep = [k for k in range(uniq.size) if new_blocks[k].size]
        return Tensor(reps[:nrr], reps[nrr:], [new_blocks[k] for k in keep], uniq[keep], sig)

    def transpose(self):
        """(symmetric_tensor.py:569-576, abelian:385-393)"""
        conj = (-self.block_irreps.astype(np.int16) + 128) % 256 - 128
        so = np.argsort(conj, kind="stable")
        s = self.signature[np.arange(-self.ndim + self.nrr, self.nrr) % self.ndim]
        return Tensor(self.col_reps, self.row_reps, [self.blocks[i].T for i in so], conj[so], s)

    def dagger(self):
        """(symmetric_tensor.py:596-609)"""
        s = np.concatenate([~self.signature[self.nrr :], ~self.signature[: self.nrr]])
        return Tensor(self.col_reps, self.row_reps, [b.T for b in self.blocks], self.block_irreps, s)

    def __matmul__(self, other):
        """Merge-join on block irreps, per-block product (symmetric_tensor.py:453-489)."""
        blocks, irreps = [], []
        oi = {int(r): i for i, r in enumerate(other.block_irreps)}
        for irr, b in zip(self.block_irreps, self.blocks, strict=True):
            if int(irr) in oi:
                blocks.append(b @ other.blocks[oi[int(irr)]])
                irreps.append(irr)
        sig = np.concatenate([self.signature[: self.nrr], other.signature[other.nrr :]])
        return Tensor(self.row_reps, other.col_reps, blocks, irreps, sig)

    def __mul__(self, x):
        return Tensor(self.row_reps, self.col_reps, [x * b for b in self.blocks], self.block_irreps, self.signature)

    def norm(self):
        return float(np.sqrt(sum(np.sum(b * b) for b in self.blocks)))

    def diag_mul(self, s_blocks, s_irreps, left=False):
        """Scale columns (or rows) by diagonal weights (symmetric_tensor.py:1189-1248)."""
        irreps = np.asarray(s_irreps, dtype=np.int16)
        if (left and self.signature[0]) or (not left and not self.signature[-1]):
            irreps = (-irreps + 128) % 256 - 128
        d = {int(r): s for r, s in zip(irreps, s_blocks, strict=True)}
        blocks, kept = [], []
        for irr, b in zip(self.block_irreps, self.blocks, strict=True):
            if int(irr) in d:
                w = d[int(irr)]
                blocks.append(b * (w[:, None] if left else w[None, :]))
                kept.append(irr)
        return Tensor(self.row_reps, self.col_reps, blocks, kept, self.signature)

    def __add__(self, other):
        """Blockwise sum keeping blocks missing on one side (symmetric_tensor.py:349-372)."""
        d = {int(r): b.copy() for r, b in zip(self.block_irreps, self.blocks, strict=True)}
        for r, b in zip(other.block_irreps, other.blocks, strict=True):
            d[int(r)] = d[int(r)] + b if int(r) in d else b.copy()
        irr = sorted(d)
        return Tensor(self.row_reps, self.col_reps, [d[r] for r in irr], irr, self.signature)

    def trace(self):
        return float(sum(np.trace(b) for b in self.blocks))

    def truncated_eigh(self, nvals, atol=0.0, rtol=None, degen_ratio=1.0):
        """Dense per-block eigh, keep the nvals largest |eigenvalues| with multiplets
        (symmetric_tensor.py:1087-1132, 1250-1434, svd_tools.py:77-113). Returns
        (value blocks, irreps, U)."""
        irr_all, shapes = self.block_sizes(self.row_reps, self.col_reps, self.signature)
        have = {int(r): b for r, b in zip(self.block_irreps, self.blocks, strict=True)}
        vals, absv, vecs = [], [], []
        for r in irr_all:
            k = nvals + 1
            if int(r) not in have:
                vals.append(np.zeros((0,)))
                absv.append(np.zeros((1,)))
                vecs.append(None)
                continue
            w, v = lg.eigh(have[int(r)], driver="evd" if have[int(r)].size > 1 else "evr")
            so = np.abs(w).argsort()[: -k - 1 : -1]
            vals.append(w[so])
            absv.append(np.abs(w[so]))
            vecs.append(v[:, so])
        cuts = find_block_cuts(absv, nvals, atol=atol, rtol=rtol, degen_ratio=degen_ratio)
        keep = [i for i in range(len(cuts)) if cuts[i] > 0 and vecs[i] is not None]
        irr = irr_all[keep]
        mid = np.repeat(irr, cuts[keep]).astype(np.int8)
        U = Tensor(self.row_reps, (mid,), [vecs[i][:, : cuts[i]] for i in keep], irr,
                   np.concatenate([self.signature[: self.nrr], [True]]))
        return [vals[i][: cuts[i]] for i in keep], irr, U

    def truncated_svd(self, cut, atol=0.0, rtol=None, degen_ratio=1.0):
        """Dense per-block SVD + find_block_cuts (symmetric_tensor.py:835-877,
        misc_tools/svd_tools.py:116-151). Returns (U, s_blocks, s_irreps, V)."""
        us, ss, vs = [], [], []
        for b in self.blocks:
            u, s, v = lg.svd(b, full_matrices=False, lapack_driver="gesdd")
            us.append(u)
            ss.append(s)
            vs.append(v)
        cuts = find_block_cuts(ss, cut, atol=atol, rtol=rtol, degen_ratio=degen_ratio)
        keep = [i for i in range(len(cuts)) if cuts[i] > 0]
        irr = self.block_irreps[keep]
        mid = np.repeat(irr, cuts[keep]).astype(np.int8)
        U = Tensor(self.row_reps, (mid,), [us[i][:, : cuts[i]] for i in keep], irr,
                   np.concatenate([self.signature[: self.nrr], [True]]))
        V = Tensor((mid,), self.col_reps, [vs[i][: cuts[i]] for i in keep], irr,
                   np.concatenate([[False], self.signature[self.nrr :]]))
        return U, [ss[i][: cuts[i]] for i in keep], irr, V
