"""Oracle: SU(2) block-sparse tensor in numpy (TEST INFRASTRUCTURE ONLY).

Representations are int arrays (2, n): row 0 = degeneracies, row 1 = irreps
labelled by dimension 2s+1. One block per irrep shared by the fused rows and
columns, sorted by irrep. Follows
frostspin/symmetric_tensor/{su2_symmetric_tensor,lie_group_symmetric_tensor,
non_abelian_symmetric_tensor}.py.

Clebsch-Gordan coefficients come from sympy (exact, as the reference's table
generator su2_clebsch_gordan.py:8-32), independently of the product's Racah
implementation.

Pinned by (tests/test_oracle_su2.py):
* clebsch_gordan: the reference's own coefficient table (fixture extracted
  from its JSON data file), orthogonality;
* combine / block_sizes: character counting (dimension sum and Sz-multiplet
  counts) on small reps;
* from_array / toarray: round trip on invariant arrays, the known-answer
  blocks of test_su2_tensor.py:24-30 and 55-68 (-0.75 / +0.25), and the cast to
  U(1) (same dense array, su2_symmetric_tensor.py:257-278);
* permute: computed through the dense array (toarray, transpose, from_array),
  which test_su2_tensor.py:36-79 asserts equals the reference's structural
  permute; norm preservation;
* basis conventions inside a block for arbitrary reps (elementary block order,
  internal/external degeneracy order, spin-reversal signs): the blocks and
  permutes the reference itself wrote for representations with external
  degeneracies and mixed signatures (tests/golden/ref_su2_layouts.npz, checked
  in tests/test_reference_fixtures.py).
"""

from functools import lru_cache

import numpy as np
import scipy.linalg as lg

from .abelian import find_block_cuts


@lru_cache(maxsize=None)
def clebsch_gordan(d1, d2, d3):
    """(d1, d2, d3) tensor, index i <-> m = s - i (su2_clebsch_gordan.py:8-32)."""
    import sympy as sp
    from sympy.physics.quantum.cg import CG

    s1, s2, s3 = (sp.Rational(d - 1, 2) for d in (d1, d2, d3))
    p = np.zeros((d1, d2, d3))
    for i1 in range(d1):
        for i2 in range(d2):
            for i3 in range(d3):
                m1, m2, m3 = s1 - i1, s2 - i2, s3 - i3
                if m1 + m2 == m3:
                    p[i1, i2, i3] = float(CG(s1, m1, s2, m2, s3, m3).doit())
    return p


def combine(reps):
    """SU(2) fusion of (degen, irrep) reps (su2_symmetric_tensor.py:109-127)."""
    degen, irreps = np.asarray(reps[0][0]), np.asarray(reps[0][1])
    for r in reps[1:]:
        acc = np.zeros((irreps[-1] + r[1][-1] - 1,), dtype=np.int64)
        for d1, i1 in zip(degen, irreps, strict=True):
            for d2, i2 in zip(r[0], r[1], strict=True):
                for irr in range(abs(i1 - i2), i1 + i2 - 1, 2):
                    acc[irr] += d1 * d2
        nz = acc.nonzero()[0]
        degen, irreps = acc[nz], nz + 1
    return np.array([degen, irreps])


def _sz_reversal_signs(d):
    return ((np.arange(d % 2, d + d % 2) % 2) * 2 - 1).astype(float)


def fusion_tensor(rep1, rep2, s1, s2, max_irrep=2**30):
    """CG projector of rep1 x rep2 onto its irreps (su2_symmetric_tensor.py:14-106)."""
    degen, irreps = combine([rep1, rep2])
    keep = irreps <= max_irrep
    degen, irreps = degen[keep], irreps[keep]
    proj = np.zeros((rep1[0] @ rep1[1], rep2[0] @ rep2[1], degen @ irreps))
    if irreps.size == 0:
        return np.zeros((2, 0), dtype=int), proj
    shifts3 = {}
    n = 0
    for d3, irr3 in zip(degen, irreps, strict=True):
        shifts3[irr3] = n
        n += d3 * irr3
    shift1 = 0
    for d1, irr1 in zip(rep1[0], rep1[1], strict=True):
        diag1 = _sz_reversal_signs(irr1)[:, None, None]
        shift2 = 0
        for d2, irr2 in zip(rep2[0], rep2[1], strict=True):
            diag2 = _sz_reversal_signs(irr2)[None, :, None]
            for irr3 in range(abs(irr1 - irr2) + 1, min(irr1 + irr2, max_irrep + 1), 2):
                p = clebsch_gordan(irr1, irr2, irr3)
                if s1:
                    p = p[::-1] * diag1
                if s2:
                    p = p[:, ::-1] * diag2
                for i3 in range(d1 * d2):
                    i1, i2 = divmod(i3, d2)
                    proj[
                        shift1 + i1 * irr1 : shift1 + (i1 + 1) * irr1,
                        shift2 + i2 * irr2 : shift2 + (i2 + 1) * irr2,
                        shifts3[irr3] + i3 * irr3 : shifts3[irr3] + (i3 + 1) * irr3,
                    ] = p
                shifts3[irr3] += d1 * d2 * irr3
            shift2 += d2 * irr2
        shift1 += d1 * irr1
    return np.array([degen, irreps]), proj


def cg_tree(reps, signature, target_irreps=None):
    """Left-to-right chained fusion (su2_symmetric_tensor.py:213-252)."""
    reps = list(reps)
    signature = list(signature)
    if len(reps) == 1:
        reps.append(np.ones((2, 1), dtype=int))
        signature.append(False)
    n = len(reps)
    if target_irreps is None:
        max_irrep = 2**30
    else:
        max_irrep = target_irreps[-1] + sum(r[1][-1] for r in reps[2:]) - n + 2
    nr, proj = fusion_tensor(reps[0], reps[1], signature[0], signature[1], max_irrep)
    if nr.size == 0:
        return nr, np.zeros((proj.shape[0] * proj.shape[1], 0))
    for i in range(2, n):
        max_irrep -= reps[i][1][-1] - 1
        nr, p = fusion_tensor(nr, reps[i], False, signature[i], max_irrep)
        if nr.size == 0:
            return nr, np.zeros((proj.size // max(p.shape[0], 1), 0))
        proj = proj.reshape(-1, p.shape[0]) @ p.reshape(p.shape[0], -1)
    return nr, proj.reshape(-1, nr[0] @ nr[1])


def sliced_elementary_trees(reps, signature, target_irreps):
    """Per elementary block, per target irrep: tree (*dims, internal, irrep)
    (lie_group_symmetric_tensor.py:365-430)."""
    per_axis = [r.shape[1] for r in reps]
    n_ele = int(np.prod(per_axis))
    internal = np.zeros((n_ele, len(target_irreps)), dtype=int)
    trees = np.empty((n_ele, len(target_irreps)), dtype=object)
    for i_ele in range(n_ele):
        mul = np.unravel_index(i_ele, per_axis)
        ele = [np.array([[1], [r[1][mul[i]]]]) for i, r in enumerate(reps)]
        rep, tree = cg_tree(ele, signature, target_irreps)
        k = 0
        sh = [r[1, 0] for r in ele] + [-1, -1]
        for i in range(rep.shape[1]):
            ideg, irr = rep[0, i], rep[1, i]
            bi = np.searchsorted(target_irreps, irr)
            if bi < len(target_irreps) and target_irreps[bi] == irr:
                sh[-2], sh[-1] = ideg, irr
                trees[i_ele, bi] = tree[:, k : k + ideg * irr].reshape(sh)
                internal[i_ele, bi] = ideg
            k += ideg * irr
    return internal, trees


def _external_degen(reps):
    per_axis = [r.shape[1] for r in reps]
    out = np.empty((int(np.prod(per_axis)), len(reps)), dtype=int)
    for i in range(out.shape[0]):
        mul = np.unravel_index(i, per_axis)
        out[i] = [r[0][mul[j]] for j, r in enumerate(reps)]
    return out


def _shifts(reps):
    return [np.concatenate([[0], np.cumsum(r[0] * r[1])]) for r in reps]


class SU2Tensor:
    def __init__(self, row_reps, col_reps, blocks, block_irreps, signature):
        self.row_reps = tuple(np.asarray(r, dtype=int) for r in row_reps)
        self.col_reps = tuple(np.asarray(r, dtype=int) for r in col_reps)
        self.blocks = tuple(np.asarray(b, dtype=np.float64) for b in blocks)
        self.block_irreps = np.asarray(block_irreps, dtype=int)
        self.signature = np.asarray(signature, dtype=bool)
        self.nrr = len(self.row_reps)
        self.shape = tuple(int(r[0] @ r[1]) for r in self.row_reps + self.col_reps)
        self.ndim = len(self.shape)

    @property
    def nblocks(self):
        return len(self.blocks)

    @staticmethod
    def block_sizes(row_reps, col_reps, signature):
        """(non_abelian_symmetric_tensor.py:36-67)"""
        rt = combine(row_reps)
        ct = combine(col_reps)
        ri, ci = (rt[1][:, None] == ct[1]).nonzero()
        return rt[1][ri], np.array([rt[0][ri], ct[0][ci]]).T.reshape(-1, 2)

    @classmethod
    def random(cls, row_reps, col_reps, signature, rng):
        irr, shapes = cls.block_sizes(row_reps, col_reps, signature)
        return cls(row_reps, col_reps, [rng.random(tuple(s)) - 0.5 for s in shapes], irr, signature)

    @classmethod
    def from_array(cls, arr, row_reps, col_reps, signature):
        """(lie_group_symmetric_tensor.py:856-967)"""
        row_reps = [np.asarray(r) for r in row_reps]
        col_reps = [np.asarray(r) for r in col_reps]
        signature = np.asarray(signature, dtype=bool)
        nrr = len(row_reps)
        ndim = nrr + len(col_reps)
        arr = np.asarray(arr, dtype=np.float64).reshape([r[0] @ r[1] for r in row_reps + col_reps])
        irreps, shapes = cls.block_sizes(row_reps, col_reps, signature)
        rsh, csh = _shifts(row_reps), _shifts(col_reps)
        rper = [r.shape[1] for r in row_reps]
        cper = [r.shape[1] for r in col_reps]
        perm = tuple(range(1, 2 * ndim, 2)) + tuple(range(0, 2 * ndim, 2))
        idr, rtrees = sliced_elementary_trees(row_reps, signature[:nrr], irreps)
        idc, ctrees = sliced_elementary_trees(col_reps, ~signature[nrr:], irreps)
        edr = _external_degen(row_reps).prod(axis=1)
        edc = _external_degen(col_reps).prod(axis=1)
        rshift = (edr[:, None] * idr).cumsum(axis=0)
        cshift = (edc[:, None] * idc).cumsum(axis=0)
        blocks = [np.empty(tuple(s)) for s in shapes]
        for ir in range(idr.shape[0]):
            rm = np.unravel_index(ir, rper)
            rdims = [row_reps[i][1][rm[i]] for i in range(nrr)]
            rdeg = [row_reps[i][0][rm[i]] for i in range(nrr)]
            rsl = [slice(rsh[i][rm[i]], rsh[i][rm[i] + 1]) for i in range(nrr)]
            for ic in range(idc.shape[0]):
                if idr[ir] @ idc[ic] == 0:
                    continue
                cm = np.unravel_index(ic, cper)
                cdims = [col_reps[i][1][cm[i]] for i in range(ndim - nrr)]
                cdeg = [col_reps[i][0][cm[i]] for i in range(ndim - nrr)]
                csl = [slice(csh[i][cm[i]], csh[i][cm[i] + 1]) for i in range(ndim - nrr)]
                sh = []
                for g, d in zip(rdeg + cdeg, rdims + cdims, strict=True):
                    sh += [g, d]
                rdim, cdim = int(np.prod(rdims)), int(np.prod(cdims))
                data = arr[tuple(rsl + csl)].reshape(sh).transpose(perm)
                data = data.reshape(rdim * cdim, edr[ir] * edc[ic])
                for bi in (idr[ir] * idc[ic]).nonzero()[0]:
                    rt = rtrees[ir, bi].reshape(-1, irreps[bi])
                    ct = ctrees[ic, bi].reshape(-1, irreps[bi])
                    proj = (rt @ ct.T).reshape(rdim, idr[ir, bi], cdim, idc[ic, bi])
                    proj = proj.swapaxes(1, 2).reshape(rdim * cdim, -1)
                    db = (proj.T @ data).reshape(idr[ir, bi], idc[ic, bi], edr[ir], edc[ic])
                    db = db.swapaxes(1, 2).reshape(idr[ir, bi] * edr[ir], idc[ic, bi] * edc[ic])
                    blocks[bi][rshift[ir, bi] - db.shape[0] : rshift[ir, bi],
                               cshift[ic, bi] - db.shape[1] : cshift[ic, bi]] = db / irreps[bi]
        return cls(row_reps, col_reps, blocks, irreps, signature)

    def toarray(self, as_matrix=False):
        """(lie_group_symmetric_tensor.py:969-1109); missing blocks read as zero."""
        nrr, ndim = self.nrr, self.ndim
        rr, cr = list(self.row_reps), list(self.col_reps)
        irreps = self.block_irreps
        rsh, csh = _shifts(rr), _shifts(cr)
        rper = [r.shape[1] for r in rr]
        cper = [r.shape[1] for r in cr]
        perm = tuple(range(1, 2 * ndim, 2)) + tuple(range(0, 2 * ndim, 2))
        inv = np.argsort(perm)
        idr, rtrees = sliced_elementary_trees(rr, self.signature[:nrr], irreps)
        idc, ctrees = sliced_elementary_trees(cr, ~self.signature[nrr:], irreps)
        edr = _external_degen(rr).prod(axis=1)
        edc = _external_degen(cr).prod(axis=1)
        rshift = (edr[:, None] * idr).cumsum(axis=0)
        cshift = (edc[:, None] * idc).cumsum(axis=0)
        arr = np.zeros(self.shape)
        for ir in range(idr.shape[0]):
            rm = np.unravel_index(ir, rper)
            rdims = [rr[i][1][rm[i]] for i in range(nrr)]
            rdeg = [rr[i][0][rm[i]] for i in range(nrr)]
            rsl = [slice(rsh[i][rm[i]], rsh[i][rm[i] + 1]) for i in range(nrr)]
            for ic in range(idc.shape[0]):
                if idr[ir] @ idc[ic] == 0:
                    continue
                cm = np.unravel_index(ic, cper)
                cdims = [cr[i][1][cm[i]] for i in range(ndim - nrr)]
                cdeg = [cr[i][0][cm[i]] for i in range(ndim - nrr)]
                csl = [slice(csh[i][cm[i]], csh[i][cm[i] + 1]) for i in range(ndim - nrr)]
                rdim, cdim = int(np.prod(rdims)), int(np.prod(cdims))
                ede = edr[ir] * edc[ic]
                dense = np.zeros((rdim * cdim, ede))
                for bi in (idr[ir] * idc[ic]).nonzero()[0]:
                    nr_, nc_ = idr[ir, bi], idc[ic, bi]
                    blk = self.blocks[bi][rshift[ir, bi] - edr[ir] * nr_ : rshift[ir, bi],
                                          cshift[ic, bi] - edc[ic] * nc_ : cshift[ic, bi]]
                    blk = blk.reshape(nr_, edr[ir], nc_, edc[ic]).swapaxes(1, 2).reshape(nr_ * nc_, ede)
                    rt = rtrees[ir, bi].reshape(-1, irreps[bi])
                    ct = ctrees[ic, bi].reshape(-1, irreps[bi])
                    proj = (rt @ ct.T).reshape(rdim, nr_, cdim, nc_).swapaxes(1, 2)
                    dense += proj.reshape(rdim * cdim, nr_ * nc_) @ blk
                ele = dense.reshape(list(rdims + cdims) + list(rdeg + cdeg)).transpose(inv)
                arr[tuple(rsl + csl)] = ele.reshape([g * d for g, d in zip(rdeg + cdeg, rdims + cdims, strict=True)])
        m = arr.reshape(int(np.prod(self.shape[:nrr])), -1)
        return m if as_matrix else arr

    def permute(self, row_axes, col_axes):
        """Dense route: from_array(toarray().transpose(axes)). test_su2_tensor.py:36-79
        asserts the reference's structural permute (lie:1116-1134) equals it."""
        axes = tuple(row_axes) + tuple(col_axes)
        reps = [(self.row_reps + self.col_reps)[a] for a in axes]
        nrr = len(row_axes)
        sig = self.signature[list(axes)]
        return SU2Tensor.from_array(self.toarray().transpose(axes), reps[:nrr], reps[nrr:], sig)

    def transpose(self):
        """Permute of rows and columns (lie_group_symmetric_tensor.py:1111-1114)."""
        n = self.ndim
        return self.permute(tuple(range(self.nrr, n)), tuple(range(self.nrr)))

    def dagger(self):
        s = np.concatenate([~self.signature[self.nrr :], ~self.signature[: self.nrr]])
        return SU2Tensor(self.col_reps, self.row_reps, [b.T for b in self.blocks], self.block_irreps, s)

    def __matmul__(self, other):
        blocks, irreps = [], []
        oi = {int(r): i for i, r in enumerate(other.block_irreps)}
        for irr, b in zip(self.block_irreps, self.blocks, strict=True):
            if int(irr) in oi:
                blocks.append(b @ other.blocks[oi[int(irr)]])
                irreps.append(irr)
        sig = np.concatenate([self.signature[: self.nrr], other.signature[other.nrr :]])
        return SU2Tensor(self.row_reps, other.col_reps, blocks, irreps, sig)

    def norm(self):
        """sqrt(sum_b dim(irrep_b) ||b||^2) (symmetric_tensor.py:624-631)."""
        return float(np.sqrt(sum(d * np.sum(b * b) for d, b in zip(self.block_irreps, self.blocks, strict=True))))

    def trace(self):
        return float(sum(d * np.trace(b) for d, b in zip(self.block_irreps, self.blocks, strict=True)))

    def truncated_svd(self, cut, atol=0.0, rtol=None, degen_ratio=1.0):
        """(symmetric_tensor.py:835-877) with irrep-dimension weights."""
        us, ss, vs = [], [], []
        for b in self.blocks:
            u, s, v = lg.svd(b, full_matrices=False, lapack_driver="gesdd")
            us.append(u)
            ss.append(s)
            vs.append(v)
        cuts = find_block_cuts(ss, cut, dims=self.block_irreps, atol=atol, rtol=rtol, degen_ratio=degen_ratio)
        keep = [i for i in range(len(cuts)) if cuts[i] > 0]
        irr = self.block_irreps[keep]
        mid = np.array([cuts[keep], irr])
        U = SU2Tensor(self.row_reps, (mid,), [us[i][:, : cuts[i]] for i in keep], irr,
                      np.concatenate([self.signature[: self.nrr], [True]]))
        V = SU2Tensor((mid,), self.col_reps, [vs[i][: cuts[i]] for i in keep], irr,
                      np.concatenate([[False], self.signature[self.nrr :]]))
        return U, [ss[i][: cuts[i]] for i in keep], irr, V
