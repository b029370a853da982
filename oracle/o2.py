"""Oracle: O(2) block tensors in numpy (TEST INFRASTRUCTURE ONLY).

Follows frostspin/symmetric_tensor/o2_symmetric_tensor.py: an O(2) tensor is
stored as the charge > 0 blocks of a U(1) tensor plus the charge-0 block split
into 0odd (-1) / 0even (0) blocks. This restatement uses explicit loops over
states (the reference's loop structure, o2:118-266, 268-352) on top of the
abelian oracle tensor.

Pinned by (tests/test_oracle_o2.py):
* fusion: Sz-weight counting of the U(1) cast;
* from_array / toarray: round trip; the dense arrays are invariant under the
  reflection (perm + sign) on every leg and conserve the U(1) charge; split /
  merge of the charge-0 block is orthogonal (norm preserved);
* permute / transpose: through the U(1) tensor, the identity the reference's
  test_o2_tensor.py:18-32 asserts, and the dense route;
* test_RVB.py:38-68 cross-symmetry: the RVB tensor cast through O(2) equals
  its U(1) form.
"""

import numpy as np

from .abelian import Tensor as U1Tensor


def rep_dim(r):
    r = np.asarray(r)
    return int((r[0] * (1 + (r[1] > 0))).sum())


def rep_to_u1(r):
    """(o2:62-89)"""
    out = []
    for d, irr in zip(r[0], r[1], strict=True):
        if irr < 1:
            out += [0] * int(d)
        else:
            out += [int(irr)] * int(d) + [-int(irr)] * int(d)
    return np.array(out, dtype=np.int8)


def reflection(r):
    """(o2:92-115)"""
    perm, sign = [], []
    k = 0
    for d, irr in zip(r[0], r[1], strict=True):
        d = int(d)
        if irr < 1:
            perm += list(range(k, k + d))
            sign += [-1 if irr == -1 else 1] * d
            k += d
        else:
            perm += list(range(k + d, k + 2 * d)) + list(range(k, k + d))
            sign += [1] * d + [-1 if irr % 2 else 1] * d
            k += 2 * d
    return np.array(perm), np.array(sign)


def combine(reps):
    """(o2:12-53)"""
    degen = {int(i): int(d) for d, i in zip(reps[0][0], reps[0][1], strict=True)}
    for r in reps[1:]:
        new = {}
        for i1, d1 in degen.items():
            for d2, i2 in zip(r[0], r[1], strict=True):
                i2, d = int(i2), d1 * int(d2)
                if i1 < 1 and i2 < 1:
                    outs = [0 if i1 == i2 else -1]
                elif i1 < 1 or i2 < 1:
                    outs = [max(i1, i2)]
                elif i1 == i2:
                    outs = [-1, 0, 2 * i1]
                else:
                    outs = [i1 + i2, abs(i1 - i2)]
                for o in outs:
                    new[o] = new.get(o, 0) + d
        degen = new
    irr = sorted(k for k, v in degen.items() if v)
    return np.array([[degen[i] for i in irr], irr])


def b0_arrays(reps, sz):
    """Odd / even rows of the charge-0 block: lists of ((col, col'), (w, w'))
    (o2:118-266), loop form."""
    shape = [rep_dim(r) for r in reps]
    refl = [reflection(r) for r in reps]
    states = [int(s) for s in np.flatnonzero(sz == 0)]
    pos = {s: i for i, s in enumerate(states)}
    odd_fx, even_fx, doublets = [], [], []
    rs, sg = [], []
    for s in states:
        idx = np.unravel_index(s, shape)
        t = 0
        sign = 1
        for k in range(len(reps)):
            t = t * shape[k] + int(refl[k][0][idx[k]])
            sign *= int(refl[k][1][idx[k]])
        rs.append(t)
        sg.append(sign)
    for i, s in enumerate(states):
        if s < rs[i]:
            doublets.append(i)
        elif s == rs[i]:
            (odd_fx if sg[i] < 0 else even_fx).append(i)
    odd = [((i, i), (1.0, 0.0)) for i in odd_fx]
    even = [((i, i), (1.0, 0.0)) for i in even_fx]
    h = 1.0 / np.sqrt(2.0)
    for i in doublets:
        j = pos[rs[i]]
        odd.append(((i, j), (h, -sg[j] * h)))
        even.append(((i, j), (h, sg[j] * h)))
    return odd, even


def _project(b0, rows, cols):
    out = np.zeros((len(rows), len(cols)))
    for a, ((r0, r1), (w0, w1)) in enumerate(rows):
        for c, ((c0, c1), (v0, v1)) in enumerate(cols):
            out[a, c] = (w0 * v0 * b0[r0, c0] + w0 * v1 * b0[r0, c1]
                         + w1 * v0 * b0[r1, c0] + w1 * v1 * b0[r1, c1])
    return out


def _unproject(b, rows, cols, b0):
    for a, ((r0, r1), (w0, w1)) in enumerate(rows):
        for c, ((c0, c1), (v0, v1)) in enumerate(cols):
            for ri, wr in ((r0, w0), (r1, w1)):
                for ci, wc in ((c0, v0), (c1, v1)):
                    b0[ri, ci] += wr * wc * b[a, c]


class O2Tensor:
    def __init__(self, row_reps, col_reps, blocks, block_irreps, signature):
        self.row_reps = tuple(np.asarray(r, dtype=int) for r in row_reps)
        self.col_reps = tuple(np.asarray(r, dtype=int) for r in col_reps)
        self.blocks = tuple(np.asarray(b, dtype=np.float64) for b in blocks)
        self.block_irreps = np.asarray(block_irreps, dtype=int)
        self.signature = np.asarray(signature, dtype=bool)
        self.nrr = len(self.row_reps)
        self.shape = tuple(rep_dim(r) for r in self.row_reps + self.col_reps)
        self.ndim = len(self.shape)

    @staticmethod
    def block_sizes(row_reps, col_reps, signature):
        rt, ct = combine(row_reps), combine(col_reps)
        ri, ci = (rt[1][:, None] == ct[1]).nonzero()
        return rt[1][ri], np.array([rt[0][ri], ct[0][ci]]).T.reshape(-1, 2)

    @classmethod
    def random(cls, row_reps, col_reps, signature, rng):
        irr, shapes = cls.block_sizes(row_reps, col_reps, signature)
        return cls(row_reps, col_reps, [rng.random(tuple(s)) - 0.5 for s in shapes], irr, signature)

    def _sz(self):
        u1r = [rep_to_u1(r) for r in self.row_reps]
        u1c = [rep_to_u1(r) for r in self.col_reps]
        from .abelian import fuse

        return u1r, u1c, fuse(u1r, self.signature[: self.nrr]), fuse(u1c, ~self.signature[self.nrr :])

    @classmethod
    def from_u1(cls, t, row_reps, col_reps):
        """(o2:637-664, 744-764)"""
        blocks, irreps = [], []
        obj = cls(row_reps, col_reps, [], [], t.signature)
        _, _, rsz, csz = obj._sz()
        for irr, b in zip(t.block_irreps, t.blocks, strict=True):
            if irr == 0:
                ro, re = b0_arrays(obj.row_reps, rsz)
                co, ce = b0_arrays(obj.col_reps, csz)
                for lab, rr, cc in ((-1, ro, co), (0, re, ce)):
                    if rr and cc:
                        blocks.append(_project(b, rr, cc))
                        irreps.append(lab)
        for irr, b in zip(t.block_irreps, t.blocks, strict=True):
            if irr > 0:
                blocks.append(b)
                irreps.append(int(irr))
        return cls(row_reps, col_reps, blocks, irreps, t.signature)

    @classmethod
    def from_array(cls, arr, row_reps, col_reps, signature):
        u1 = U1Tensor.from_array(arr, [rep_to_u1(r) for r in row_reps], [rep_to_u1(r) for r in col_reps], signature)
        return cls.from_u1(u1, row_reps, col_reps)

    def to_u1(self):
        """(o2:672-728, 798-851)"""
        u1r, u1c, rsz, csz = self._sz()
        rshape = [rep_dim(r) for r in self.row_reps]
        cshape = [rep_dim(r) for r in self.col_reps]
        rrefl = [reflection(r) for r in self.row_reps]
        crefl = [reflection(r) for r in self.col_reps]

        def mirrored(states, shape, refl):
            out, sign = [], []
            for s in states:
                idx = np.unravel_index(s, shape)
                t, sg = 0, 1
                for k in range(len(shape)):
                    t = t * shape[k] + int(refl[k][0][idx[k]])
                    sg *= int(refl[k][1][idx[k]])
                out.append(t)
                sign.append(sg)
            return np.array(out), np.array(sign)

        blocks, irreps = {}, []
        for irr, b in zip(self.block_irreps, self.blocks, strict=True):
            if irr > 0:
                blocks[int(irr)] = b
                rm, rsg = mirrored(np.flatnonzero(rsz == irr), rshape, rrefl)
                cm, csg = mirrored(np.flatnonzero(csz == irr), cshape, crefl)
                neg = np.empty(b.shape)
                ro = rm.argsort().argsort()
                co = cm.argsort().argsort()
                for i in range(b.shape[0]):
                    for j in range(b.shape[1]):
                        neg[ro[i], co[j]] = rsg[i] * csg[j] * b[i, j]
                blocks[-int(irr)] = neg
        zero = {int(i): b for i, b in zip(self.block_irreps, self.blocks, strict=True) if i < 1}
        if zero:
            ro, re = b0_arrays(self.row_reps, rsz)
            co, ce = b0_arrays(self.col_reps, csz)
            b0 = np.zeros((int((rsz == 0).sum()), int((csz == 0).sum())))
            if -1 in zero:
                _unproject(zero[-1], ro, co, b0)
            if 0 in zero:
                _unproject(zero[0], re, ce, b0)
            blocks[0] = b0
        irr = sorted(blocks)
        return U1Tensor(u1r, u1c, [blocks[i] for i in irr], np.array(irr, dtype=np.int8), self.signature)

    def toarray(self):
        return self.to_u1().toarray()

    def permute(self, row_axes, col_axes):
        """Through the U(1) tensor (test_o2_tensor.py:18-32)."""
        axes = tuple(row_axes) + tuple(col_axes)
        reps = [(self.row_reps + self.col_reps)[a] for a in axes]
        tp = self.to_u1().permute(row_axes, col_axes)
        return O2Tensor.from_u1(tp, reps[: len(row_axes)], reps[len(row_axes) :])

    def norm(self):
        return float(np.sqrt(sum((1 + (i > 0)) * np.sum(b * b) for i, b in zip(self.block_irreps, self.blocks, strict=True))))
