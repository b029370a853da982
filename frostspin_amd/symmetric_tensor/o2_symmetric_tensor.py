"""O(2) symmetric tensors (o2_symmetric_tensor.py:583-945).

O(2) = U(1) x| Z2. Irreps: -1 = 0odd, 0 = 0even (dimension 1), n > 0 the
doublet (+n, -n). A representation is an int array (2, k) of (degeneracies,
irreps), irreps sorted. The basis is the U(1) one (o2:62-89): a leg lists its
0odd states, its 0even states, then for each n > 0 the +n states followed by
the -n states. An O(2) tensor stores

* for n > 0, the U(1) block of charge +n (the -n block is its reflection, with
  the permutation and signs of o2:92-115);
* for the charge-0 U(1) block, its split into 0odd and 0even blocks: reflection
  fixed points keep weight 1, every doublet of states gives one odd and one
  even combination with weights +-1/sqrt(2) (o2:118-266).

Every data movement runs on the device: the Sz = 0 split / merge and the
reflected blocks are two-term weighted gathers (fs_batched_gather2d_f64); the
permute goes through the U(1) tensor and its planned gather, exactly the map
the reference's own tests assert (test_o2_tensor.py:18-32).
"""

from functools import lru_cache

import numpy as np

from .. import _ops
from .non_abelian_symmetric_tensor import NonAbelianSymmetricTensor
from .symmetric_tensor import _packed_offsets
from .tools import check_norm, symmetric_tensor_types
from .u1_symmetric_tensor import U1SymmetricTensor

_INV_SQRT2 = 1.0 / np.sqrt(2.0)


def _rep_key(rep):
    r = np.asarray(rep, dtype=np.int64)
    return tuple(r[0].tolist()), tuple(r[1].tolist())


def o2_rep_to_u1(rep):
    """int8 U(1) charges of an O(2) representation (o2:62-89)."""
    r = np.asarray(rep, dtype=np.int64)
    parts = []
    for d, irr in zip(r[0], r[1], strict=True):
        if irr < 1:
            parts.append(np.zeros((d,), dtype=np.int8))
        else:
            parts.append(np.full((d,), irr, dtype=np.int8))
            parts.append(np.full((d,), -irr, dtype=np.int8))
    return np.concatenate(parts) if parts else np.zeros((0,), dtype=np.int8)


@lru_cache(maxsize=None)
def _reflection(rep_key):
    """Permutation and sign of the basis under the reflection (o2:92-115):
    0odd states change sign, +n <-> -n swap, with a -1 on the -n -> +n image
    for odd n."""
    degen, irreps = (np.array(x, dtype=np.int64) for x in rep_key)
    dim = int((degen * (1 + (irreps > 0))).sum())
    perm = np.arange(dim, dtype=np.int64)
    sign = np.ones((dim,), dtype=np.int64)
    k = 0
    for d, irr in zip(degen, irreps, strict=True):
        d = int(d)
        if irr < 1:
            if irr == -1:
                sign[k : k + d] = -1
            k += d
            continue
        perm[k : k + d] = np.arange(k + d, k + 2 * d)
        perm[k + d : k + 2 * d] = np.arange(k, k + d)
        if irr % 2:
            sign[k + d : k + 2 * d] = -1
        k += 2 * d
    perm.flags.writeable = False
    sign.flags.writeable = False
    return perm, sign


def _fuse_pair(r1, r2):
    """O(2) product of two representations (o2:21-53)."""
    nmax = max(int(r1[1][-1]), 0) + max(int(r2[1][-1]), 0)
    degen = np.zeros((nmax + 2,), dtype=np.int64)
    for d1, i1 in zip(r1[0], r1[1], strict=True):
        for d2, i2 in zip(r2[0], r2[1], strict=True):
            d = int(d1) * int(d2)
            if i1 < 1 and i2 < 1:
                degen[(i1 + i2 + 1) % 2] += d  # 0o x 0o = 0e x 0e = 0e, mixed = 0o
            elif i1 < 1:
                degen[i2 + 1] += d
            elif i2 < 1:
                degen[i1 + 1] += d
            elif i1 == i2:
                degen[0] += d
                degen[1] += d
                degen[2 * i1 + 1] += d
            else:
                degen[i1 + i2 + 1] += d
                degen[abs(i1 - i2) + 1] += d
    irreps = np.arange(-1, nmax + 1)
    nz = degen.nonzero()[0]
    return np.array([degen[nz], irreps[nz]], dtype=np.int64)


def _sz0_maps(reps, signature):
    """Sz = 0 split of one side (o2:118-266): rows of the 0odd and 0even blocks
    as (2 column indices into the U(1) charge-0 block, 2 weights)."""
    key = (tuple(_rep_key(r) for r in reps), tuple(bool(s) for s in signature))
    return _sz0_maps_cached(key)


@lru_cache(maxsize=4096)
def _sz0_maps_cached(key):
    rep_keys, signature = key
    reps = [np.array(k, dtype=np.int64) for k in rep_keys]
    sz = U1SymmetricTensor.combine_representations([o2_rep_to_u1(r) for r in reps], np.array(signature))
    states = np.flatnonzero(sz == 0)
    shape = tuple(int((r[0] * (1 + (r[1] > 0))).sum()) for r in reps)
    multi = np.unravel_index(states, shape)
    refl_multi = []
    sign = np.ones(states.shape, dtype=np.int64)
    for k, rk in enumerate(rep_keys):
        perm, sg = _reflection(rk)
        refl_multi.append(perm[multi[k]])
        sign *= sg[multi[k]]
    refl = np.ravel_multi_index(tuple(refl_multi), shape) if states.size else states
    fixed = states == refl
    odd_fixed = np.flatnonzero(fixed & (sign < 0))
    even_fixed = np.flatnonzero(fixed & (sign > 0))
    first = np.flatnonzero(states < refl)  # doublets, first member
    partner = np.searchsorted(states, refl[first])
    nd = first.size
    o_idx = np.empty((odd_fixed.size + nd, 2), dtype=np.int64)
    o_w = np.empty((odd_fixed.size + nd, 2))
    e_idx = np.empty((even_fixed.size + nd, 2), dtype=np.int64)
    e_w = np.empty((even_fixed.size + nd, 2))
    o_idx[: odd_fixed.size] = odd_fixed[:, None]
    o_w[: odd_fixed.size] = (1.0, 0.0)
    e_idx[: even_fixed.size] = even_fixed[:, None]
    e_w[: even_fixed.size] = (1.0, 0.0)
    o_idx[odd_fixed.size :, 0] = first
    o_idx[odd_fixed.size :, 1] = partner
    o_w[odd_fixed.size :, 0] = _INV_SQRT2
    o_w[odd_fixed.size :, 1] = -sign[partner] * _INV_SQRT2
    e_idx[even_fixed.size :, 0] = first
    e_idx[even_fixed.size :, 1] = partner
    e_w[even_fixed.size :, 0] = _INV_SQRT2
    e_w[even_fixed.size :, 1] = sign[partner] * _INV_SQRT2
    n0 = states.size
    # inverse maps (merge): each charge-0 state is in at most one odd and one even row
    inv = []
    for idx, w in ((o_idx, o_w), (e_idx, e_w)):
        ii = np.zeros((n0, 2), dtype=np.int64)
        ww = np.zeros((n0, 2))
        for a in range(2):
            live = w[:, a] != 0.0
            ii[idx[live, a], 0] = np.flatnonzero(live)
            ww[idx[live, a], 0] = w[live, a]
        inv.append((ii, ww))
    out = (n0, (o_idx, o_w), (e_idx, e_w), inv[0], inv[1])
    for arr in (o_idx, o_w, e_idx, e_w, *inv[0], *inv[1]):
        arr.flags.writeable = False
    return out


def _reflected_rows(reps, signature, n):
    """Rows of the U(1) charge -n block as a one-term gather from the +n block:
    (source row index, sign) per destination row (o2:798-851)."""
    key = (tuple(_rep_key(r) for r in reps), tuple(bool(s) for s in signature), int(n))
    return _reflected_rows_cached(key)


@lru_cache(maxsize=8192)
def _reflected_rows_cached(key):
    rep_keys, signature, n = key
    reps = [np.array(k, dtype=np.int64) for k in rep_keys]
    sz = U1SymmetricTensor.combine_representations([o2_rep_to_u1(r) for r in reps], np.array(signature))
    plus = np.flatnonzero(sz == n)
    shape = tuple(int((r[0] * (1 + (r[1] > 0))).sum()) for r in reps)
    multi = np.unravel_index(plus, shape)
    refl_multi = []
    sign = np.ones(plus.shape, dtype=np.int64)
    for k, rk in enumerate(rep_keys):
        perm, sg = _reflection(rk)
        refl_multi.append(perm[multi[k]])
        sign *= sg[multi[k]]
    refl = np.ravel_multi_index(tuple(refl_multi), shape)
    order = np.argsort(refl, kind="stable")  # -n rows in stable order <- source rows
    src = order.astype(np.int64)
    w = sign[order].astype(np.float64)
    src.flags.writeable = False
    w.flags.writeable = False
    return src, w


class O2SymmetricTensor(NonAbelianSymmetricTensor):
    """SymmetricTensor with global O(2) symmetry (irreps -1 = 0odd, 0 = 0even,
    n > 0 = doublet (+n, -n))."""

    _symmetry = "O2"

    @staticmethod
    def singlet():
        return np.array([[1], [0]], dtype=np.int64)

    @staticmethod
    def representation_dimension(rep):
        r = np.asarray(rep)
        return int((r[0] * (1 + (r[1] > 0))).sum())

    @staticmethod
    def irrep_dimension(irrep):
        return 1 + int(irrep > 0)

    @staticmethod
    def combine_representations(reps, _signature):
        reps = [np.asarray(r, dtype=np.int64) for r in reps]
        out = reps[0]
        for r in reps[1:]:
            out = _fuse_pair(out, r)
        return out

    @staticmethod
    def conjugate_irrep(irr):
        return irr

    @staticmethod
    def conjugate_representation(rep):
        return rep

    # casts ------------------------------------------------------------------
    def toO2(self):
        return self

    def toabelian(self):
        return self.toU1()

    @classmethod
    def from_U1(cls, tu1, row_reps, col_reps):
        """O(2) tensor from a U(1) tensor with O(2) symmetry whose charges follow
        the O(2) basis order (o2:637-664); only its charge >= 0 blocks are read."""
        row_reps = tuple(np.asarray(r, dtype=np.int64) for r in row_reps)
        col_reps = tuple(np.asarray(r, dtype=np.int64) for r in col_reps)
        assert tu1.n_row_reps == len(row_reps) and len(tu1.col_reps) == len(col_reps)
        data, shapes, irreps = cls._blocks_from_U1(tu1, row_reps, col_reps)
        offs, _ = _packed_offsets(shapes)
        return cls._build(row_reps, col_reps, data, shapes, offs, irreps, tu1.signature)

    @classmethod
    def _blocks_from_U1(cls, tu1, row_reps, col_reps):
        """(device buffer, shapes, irreps) of the O(2) blocks of tu1 (o2:744-764)."""
        nrr = len(row_reps)
        sig = tu1.signature
        u1_irreps = np.asarray(tu1.block_irreps, dtype=np.int64)
        pos = u1_irreps[u1_irreps > 0]
        pos_ids = np.flatnonzero(u1_irreps > 0)
        zero = np.flatnonzero(u1_irreps == 0)
        plan = []  # (kind, u1 block, rows, cols)
        if zero.size:
            b = int(zero[0])
            n0r, ro, re, _, _ = _sz0_maps(row_reps, sig[:nrr])
            n0c, co, ce, _, _ = _sz0_maps(col_reps, ~sig[nrr:])
            assert tuple(tu1._bshapes[b]) == (n0r, n0c)
            for irr, rmap, cmap in ((-1, ro, co), (0, re, ce)):
                if rmap[0].shape[0] and cmap[0].shape[0]:
                    plan.append(("split", b, irr, rmap, cmap))
        for irr, b in zip(pos, pos_ids, strict=True):
            plan.append(("copy", int(b), int(irr), None, None))
        shapes = np.array(
            [(p[3][0].shape[0], p[4][0].shape[0]) if p[0] == "split" else tuple(tu1._bshapes[p[1]])
             for p in plan],
            dtype=np.int64,
        ).reshape(-1, 2)
        irreps = np.array([p[2] for p in plan], dtype=np.int64)
        offs, total = _packed_offsets(shapes)
        out = _ops.empty(total)
        copies, gathers, idx, w = [], [], [], []
        nidx = 0
        for k, (kind, b, _irr, rmap, cmap) in enumerate(plan):
            m, n = shapes[k]
            rs, cs = tu1._strides(b)
            if kind == "copy":
                copies.append((m, n, tu1._boffs[b], rs, cs, offs[k], n, 1, 0, 0, 0, 1.0, 0.0))
                continue
            sr, sc = tu1._bshapes[b]
            gathers.append((m, n, sr, sc, tu1._boffs[b], rs, cs, offs[k], n, 1, nidx, nidx,
                            nidx + 2 * m, nidx + 2 * m, 0.0))
            idx += [rmap[0].reshape(-1), cmap[0].reshape(-1)]
            w += [rmap[1].reshape(-1), cmap[1].reshape(-1)]
            nidx += 2 * (m + n)
        _ops.copy2d(copies, tu1._data, out)
        if gathers:
            _ops.gather2d(gathers, np.concatenate(idx), np.concatenate(w), tu1._data, out)
        return out, shapes, irreps

    def toU1(self):
        """U(1) tensor with the same dense array (o2:672-728): -n blocks by
        reflection, the charge-0 block merged from 0odd / 0even, +n blocks copied."""
        nrr = self._nrr
        sig = self._signature
        u1_rows = tuple(o2_rep_to_u1(r) for r in self._row_reps)
        u1_cols = tuple(o2_rep_to_u1(r) for r in self._col_reps)
        irreps = np.asarray(self._block_irreps, dtype=np.int64)
        pos_ids = np.flatnonzero(irreps > 0)
        zero_ids = {int(irreps[b]): int(b) for b in np.flatnonzero(irreps < 1)}
        blocks = []  # (u1 irrep, kind, o2 block)
        for b in pos_ids[::-1]:
            blocks.append((-int(irreps[b]), "reflect", int(b)))
        if zero_ids:
            blocks.append((0, "merge", None))
        for b in pos_ids:
            blocks.append((int(irreps[b]), "copy", int(b)))
        n0r = n0c = 0
        if zero_ids:
            n0r, _, _, rinv_o, rinv_e = _sz0_maps(self._row_reps, sig[:nrr])
            n0c, _, _, cinv_o, cinv_e = _sz0_maps(self._col_reps, ~sig[nrr:])
        shapes = np.array(
            [(n0r, n0c) if kind == "merge" else tuple(self._bshapes[b]) for _, kind, b in blocks],
            dtype=np.int64,
        ).reshape(-1, 2)
        offs, total = _packed_offsets(shapes)
        out = _ops.zeros(total) if zero_ids else _ops.empty(total)
        copies, gathers, idx, w = [], [], [], []
        nidx = 0
        for k, (u1_irr, kind, b) in enumerate(blocks):
            m, n = shapes[k]
            if kind == "copy":
                rs, cs = self._strides(b)
                copies.append((m, n, self._boffs[b], rs, cs, offs[k], n, 1, 0, 0, 0, 1.0, 0.0))
                continue
            if kind == "reflect":
                rsrc, rw = _reflected_rows(self._row_reps, sig[:nrr], -u1_irr)
                csrc, cw = _reflected_rows(self._col_reps, ~sig[nrr:], -u1_irr)
                parts = [(b, rsrc, rw, csrc, cw)]
            else:
                parts = []
                for irr, rinv, cinv in ((-1, rinv_o, cinv_o), (0, rinv_e, cinv_e)):
                    if irr in zero_ids:
                        parts.append((zero_ids[irr], rinv[0][:, 0], rinv[1][:, 0], cinv[0][:, 0], cinv[1][:, 0]))
            for src_b, ri, rw_, ci, cw_ in parts:
                rs, cs = self._strides(src_b)
                sr, sc = self._bshapes[src_b]
                gathers.append((m, n, sr, sc, self._boffs[src_b], rs, cs, offs[k], n, 1,
                                nidx, nidx, nidx + 2 * m, nidx + 2 * m, 0.0))
                ii = np.zeros((m, 2), dtype=np.int64)
                ww = np.zeros((m, 2))
                ii[:, 0] = ri
                ww[:, 0] = rw_
                jj = np.zeros((n, 2), dtype=np.int64)
                wc = np.zeros((n, 2))
                jj[:, 0] = ci
                wc[:, 0] = cw_
                idx += [ii.reshape(-1), jj.reshape(-1)]
                w += [ww.reshape(-1), wc.reshape(-1)]
                nidx += 2 * (m + n)
        _ops.copy2d(copies, self._data, out)
        # the merge writes the charge-0 block as a sum of two gathers: accumulate
        merged = {}
        final = []
        for g in gathers:
            key = g[7]
            if key in merged:
                g = (*g[:14], 1.0)
            merged[key] = True
            final.append(g)
        # gathers accumulating into the same block must not run in the same launch
        first = [g for g in final if g[14] == 0.0]
        second = [g for g in final if g[14] != 0.0]
        idx_all = np.concatenate(idx) if idx else np.zeros((0,), np.int64)
        w_all = np.concatenate(w) if w else np.zeros((0,))
        if first:
            _ops.gather2d(first, idx_all, w_all, self._data, out)
        if second:
            _ops.gather2d(second, idx_all, w_all, self._data, out)
        tu1 = U1SymmetricTensor._build(
            u1_rows, u1_cols, out, shapes, offs, np.array([x[0] for x in blocks], dtype=np.int8), sig
        )
        assert check_norm(self, tu1)
        return tu1

    # symmetry hooks -----------------------------------------------------------
    @classmethod
    def _blocks_from_dense(cls, arr, row_reps, col_reps, signature):
        """Through the U(1) tensor, then the charge-0 split (o2:736-742)."""
        row_reps = tuple(np.asarray(r, dtype=np.int64) for r in row_reps)
        col_reps = tuple(np.asarray(r, dtype=np.int64) for r in col_reps)
        tu1 = U1SymmetricTensor.from_array(
            arr, [o2_rep_to_u1(r) for r in row_reps], [o2_rep_to_u1(r) for r in col_reps], signature
        )
        return cls._blocks_from_U1(tu1, row_reps, col_reps)

    def _tomatrix(self):
        return self.toU1()._tomatrix()

    def _permute_data(self, axes, nrr):
        """Permute of the U(1) tensor, then its charge >= 0 blocks back to O(2)
        (the map test_o2_tensor.py:18-32 asserts; o2:853-936)."""
        tu1 = self.toU1()
        data, shapes, irreps = tu1._permute_data(axes, nrr)
        reps = [(self._row_reps + self._col_reps)[a] for a in axes]
        u1_reps = [(tu1.row_reps + tu1.col_reps)[a] for a in axes]
        sig = self._signature[list(axes)]
        offs, _ = _packed_offsets(shapes)
        tp = U1SymmetricTensor._build(u1_reps[:nrr], u1_reps[nrr:], data, shapes, offs, irreps, sig)
        return self._blocks_from_U1(tp, tuple(reps[:nrr]), tuple(reps[nrr:]))

    def transpose(self):
        """Rows and columns swapped: O(2) blocks of the transposed U(1) tensor
        (o2:938-942)."""
        tu1 = self.toU1().transpose()
        data, shapes, irreps = self._blocks_from_U1(tu1, self._col_reps, self._row_reps)
        offs, _ = _packed_offsets(shapes)
        return type(self)._build(self._col_reps, self._row_reps, data, shapes, offs, irreps, tu1.signature)

    def update_signature(self, sign_update):
        """Every O(2) irrep is self-conjugate and the blocks are the U(1) ones:
        only the signature changes (o2:730-741)."""
        up = np.asarray(sign_update, dtype=bool)
        assert up.shape == (self._ndim,)
        self._signature = self._signature ^ up
        assert self.check_blocks_fit_representations()


symmetric_tensor_types["O2"] = O2SymmetricTensor
