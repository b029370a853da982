"""Lie-group (SU(2)) block tensors: structural data on the host, data movement
on the device.

Storage follows lie_group_symmetric_tensor.py:856-1109: the rows of block J
are the concatenation, over the row elementary blocks (one irrep per row leg,
row-major over the legs' irrep indices), of (internal degeneracy x external
degeneracy) sub-rows, internal index major; same for columns. The internal
index counts the fusion paths of the left-to-right Clebsch-Gordan tree
(su2_symmetric_tensor.py:213-252), the external index is row-major over the
legs' degeneracy indices.

Every data movement (permute, transpose, dense <-> blocks) is one native
structural-transform plan (fs_lie_plan_*): per elementary block, a gather of
strided rows, a small dense matrix (CG overlap unitary or projector) and a
scatter. The host side computes those matrices once per set of irreps - a
simple block's unitary depends only on its irreps, the signature and the
permutation, never on degeneracies (the reference caches it the same way,
lie:540-602) - and assembles per call the offset tables with vectorized
numpy.
"""

import threading
import time
from collections import OrderedDict
from functools import reduce

import numpy as np

from .. import _lib, _ops
from . import plan_store
from .abelian_symmetric_tensor import _dense_source
from .non_abelian_symmetric_tensor import NonAbelianSymmetricTensor
from .tools import config as _debug_config

# Overlap coefficients below this magnitude are exact zeros polluted by
# rounding; they are cleared as in overlap_cg_trees (lie:720-726).
_OVERLAP_ZERO = 1e-14


def _grid(nper):
    """(prod(nper), len(nper)) row-major multi-indices."""
    return np.indices(tuple(int(n) for n in nper), dtype=np.int64).reshape(len(nper), -1).T.copy()


def _outer(vectors):
    """Row-major flattened outer product of 1-D integer vectors."""
    if not vectors:
        return np.ones((1,), dtype=np.int64)
    return reduce(np.multiply.outer, [np.asarray(v, dtype=np.int64) for v in vectors]).reshape(-1)


def _exclusive_cumsum0(a):
    out = np.zeros_like(a)
    if a.shape[0] > 1:
        out[1:] = np.cumsum(a[:-1], axis=0)
    return out


def _suffix_products(mat):
    """out[:, j] = prod(mat[:, j+1:]) (row-major strides of a mixed radix)."""
    out = np.ones_like(mat)
    for j in range(mat.shape[1] - 2, -1, -1):
        out[:, j] = out[:, j + 1] * mat[:, j + 1]
    return out


def _pad8(n):
    return (np.asarray(n, dtype=np.int64) + 7) // 8 * 8


class _PermuteStructure:
    """Irreps-only data of one permute pattern (host tables + unitaries)."""

    __slots__ = (
        "W", "idc_in", "idc_out", "idr_in", "idr_out", "ic_in", "ic_out", "in_a", "in_b",
        "in_t", "in_targets", "ir_in", "ir_out", "n_in", "n_out", "n_out_pad", "out_a",
        "out_b", "out_t", "out_targets", "sig_idx", "w_off", "pattern", "weights", "_keep",
    )

    # arrays that define a structure (persisted by plan_store)
    _ARRAYS = (
        "W", "idc_in", "idc_out", "idr_in", "idr_out", "ic_in", "ic_out", "in_a", "in_b", "in_t",
        "in_targets", "ir_in", "ir_out", "n_in", "n_out", "n_out_pad", "out_a", "out_b", "out_t",
        "out_targets", "sig_idx", "w_off",
    )

    def to_arrays(self):
        return {k: np.asarray(getattr(self, k)) for k in self._ARRAYS}

    @classmethod
    def from_arrays(cls, arrays):
        st = cls()
        for k in cls._ARRAYS:
            setattr(st, k, arrays[k])
        return st

    def bind(self, ndim, nrr, nrr_out, axes, nirr):
        """The C pattern (fs_lie_pattern) over this structure's arrays; the
        arrays are kept alive with it."""
        arrs = {
            "axes": np.asarray(axes, dtype=np.int64), "nirr": np.asarray(nirr, dtype=np.int64),
            "sig_idx": self.sig_idx, "ir_in": self.ir_in, "ic_in": self.ic_in, "ir_out": self.ir_out,
            "ic_out": self.ic_out, "n_in": self.n_in, "n_out": self.n_out, "w_off": self.w_off,
            "n_out_pad": self.n_out_pad, "idr_in": self.idr_in, "idc_in": self.idc_in,
            "idr_out": self.idr_out, "idc_out": self.idc_out, "in_t": self.in_t, "in_a": self.in_a,
            "in_b": self.in_b, "out_t": self.out_t, "out_a": self.out_a, "out_b": self.out_b,
        }
        arrs = {k: np.ascontiguousarray(v, dtype=np.int64) for k, v in arrs.items()}
        self._keep = arrs
        self.pattern = _lib.LiePattern(
            ndim, nrr, nrr_out, len(self.n_in), len(self.in_targets), len(self.out_targets),
            *(arrs[f].ctypes.data_as(_lib.c_i64p) for f, _ in _lib.LiePattern._fields_[6:]),
        )
        self.weights = _ops.LieWeights(self.pattern, self.W)


class _DenseStructure:
    """Irreps-only data of a dense <-> blocks transform."""

    __slots__ = ("ic", "idc", "idr", "ir", "m_idx", "n_rows", "proj", "sec_a", "sec_b", "sec_t",
                 "sig_idx", "targets")


class LieGroupSymmetricTensor(NonAbelianSymmetricTensor):
    """Block tensor whose structure is defined by Clebsch-Gordan fusion trees.
    Subclasses provide _elementary_trees (group-specific)."""

    _unitary_cache = {}
    _permute_structures = {}
    _structures_by_bytes = {}
    _dense_structures = {}
    _side_cache = {}
    _plans = OrderedDict()
    _plans_lock = threading.Lock()  # requests served from several host threads share the cache
    _PLAN_CACHE_SIZE = 512  # a few concurrent requests of different chi keep their plans

    # group hook ---------------------------------------------------------------
    @classmethod
    def _elementary_trees(cls, irreps, flags):
        """{irrep J: tree (prod(dim irreps), internal degeneracy, dim J)} for one
        elementary block (one irrep per leg); flags[k] applies the leg's
        signature convention."""
        raise NotImplementedError("Must be defined in derived class")

    # structural helpers -------------------------------------------------------
    @classmethod
    def _allowed_irreps(cls, irreps_per_leg, signature, nrr):
        reps = [cls.init_representation(np.ones((len(ir),), dtype=np.int64), ir) for ir in irreps_per_leg]
        block_irreps, _ = cls.get_block_sizes(reps[:nrr], reps[nrr:], np.asarray(signature, dtype=bool))
        return tuple(int(x) for x in block_irreps)

    @classmethod
    def _side_internal(cls, irreps_per_leg, flags, targets):
        """(n_elementary, n_targets) internal degeneracies of one side."""
        key = (cls, irreps_per_leg, flags, targets)
        out = cls._side_cache.get(key)
        if out is None:
            grid = _grid([len(ir) for ir in irreps_per_leg])
            out = np.zeros((grid.shape[0], len(targets)), dtype=np.int64)
            tpos = {t: i for i, t in enumerate(targets)}
            for e in range(grid.shape[0]):
                dims = tuple(irreps_per_leg[k][grid[e, k]] for k in range(grid.shape[1]))
                for J, tree in cls._elementary_trees(dims, flags).items():
                    if J in tpos:
                        out[e, tpos[J]] = tree.shape[1]
            out.flags.writeable = False
            cls._side_cache[key] = out
        return out

    @classmethod
    def _sigma_projector(cls, irreps, signature, nrr, targets=None):
        """Columns (J, a, b) of the CG projector of one elementary block:
        P[m, (J,a,b)] = sum_mu rtree[m_r, a, mu] ctree[m_c, b, mu], rows the
        row-major m multi-index over all legs. Returns (sections, P) with
        sections = [(J, n_row_internal, n_col_internal)] in ascending J."""
        rows = cls._elementary_trees(irreps[:nrr], tuple(bool(s) for s in signature[:nrr]))
        cols = cls._elementary_trees(irreps[nrr:], tuple(not s for s in signature[nrr:]))
        shared = sorted(set(rows) & set(cols))
        if targets is not None:
            shared = [J for J in shared if J in targets]
        secs, mats = [], []
        for J in shared:
            r, c = rows[J], cols[J]
            p = np.einsum("aim,bjm->abij", r, c, optimize=True)
            mats.append(p.reshape(r.shape[0] * c.shape[0], r.shape[1] * c.shape[1]))
            secs.append((J, r.shape[1], c.shape[1]))
        n = int(np.prod([cls.irrep_dimension(d) for d in irreps]))
        P = np.concatenate(mats, axis=1) if mats else np.zeros((n, 0))
        return secs, P

    @classmethod
    def _sigma_unitary(cls, irreps, signature, nrr, axes, nrr_out):
        """Overlap of the old and new CG projectors of one simple block
        (overlap_cg_trees, lie:657-733): U[(J',a',b'), (J,a,b)] =
        <P'_(J',a',b') | permuted P_(J,a,b)> / dim J'."""
        key = (cls, irreps, signature, nrr, axes, nrr_out)
        hit = cls._unitary_cache.get(key)
        if hit is not None:
            return hit
        ndim = len(irreps)
        in_secs, P = cls._sigma_projector(irreps, signature, nrr)
        dims = [cls.irrep_dimension(d) for d in irreps]
        P = P.reshape(*dims, P.shape[1]).transpose(*axes, ndim).reshape(-1, P.shape[1])
        new_irreps = tuple(irreps[a] for a in axes)
        new_sig = tuple(signature[a] for a in axes)
        out_secs, Q = cls._sigma_projector(new_irreps, new_sig, nrr_out)
        live = np.flatnonzero(np.any(P != 0.0, axis=1) & np.any(Q != 0.0, axis=1))
        U = Q[live].T @ P[live]
        U[np.abs(U) < _OVERLAP_ZERO] = 0.0
        k = 0
        for J, nr, nc in out_secs:
            U[k : k + nr * nc] /= cls.irrep_dimension(J)
            k += nr * nc
        assert U.shape[0] == U.shape[1], "number of coefficients not conserved"
        U.flags.writeable = False
        hit = (tuple(in_secs), tuple(out_secs), U)
        cls._unitary_cache[key] = hit
        return hit

    # permute --------------------------------------------------------------------
    @classmethod
    def _permute_structure(cls, irreps_per_leg, signature, nrr, axes, nrr_out):
        key = (cls, irreps_per_leg, signature, nrr, axes, nrr_out)
        st = cls._permute_structures.get(key)
        if st is not None:
            return st
        nper = np.array([len(ir) for ir in irreps_per_leg], dtype=np.int64)
        skey = plan_store.pattern_key(cls.symmetry(), irreps_per_leg, signature, nrr, axes, nrr_out)
        stored = plan_store.load(skey)
        if stored is not None:
            st = _PermuteStructure.from_arrays(stored)
            st.bind(len(irreps_per_leg), nrr, nrr_out, axes, nper)
            cls._permute_structures[key] = st
            return st
        t_build = time.perf_counter()
        ndim = len(irreps_per_leg)
        in_targets = cls._allowed_irreps(irreps_per_leg, signature, nrr)
        new_irreps = tuple(irreps_per_leg[a] for a in axes)
        new_sig = tuple(signature[a] for a in axes)
        out_targets = cls._allowed_irreps(new_irreps, new_sig, nrr_out)
        if not in_targets:
            raise ValueError("Representations do not allow any block")
        nper = np.array([len(ir) for ir in irreps_per_leg], dtype=np.int64)
        grid = _grid(nper)
        ax = np.asarray(axes)

        def ravel(cols):
            if len(cols) == 0:
                return np.zeros((grid.shape[0],), dtype=np.int64)
            return np.ravel_multi_index(tuple(grid[:, c] for c in cols), tuple(nper[list(cols)]))

        ir_in = ravel(list(range(nrr)))
        ic_in = ravel(list(range(nrr, ndim)))
        ir_out = ravel(list(ax[:nrr_out]))
        ic_out = ravel(list(ax[nrr_out:]))
        f_ri = tuple(bool(s) for s in signature[:nrr])
        f_ci = tuple(not s for s in signature[nrr:])
        f_ro = tuple(bool(s) for s in new_sig[:nrr_out])
        f_co = tuple(not s for s in new_sig[nrr_out:])
        idr_in = cls._side_internal(irreps_per_leg[:nrr], f_ri, in_targets)
        idc_in = cls._side_internal(irreps_per_leg[nrr:], f_ci, in_targets)
        idr_out = cls._side_internal(new_irreps[:nrr_out], f_ro, out_targets)
        idc_out = cls._side_internal(new_irreps[nrr_out:], f_co, out_targets)
        valid = np.flatnonzero((idr_in[ir_in] * idc_in[ic_in]).sum(axis=1) > 0)
        tin = {t: i for i, t in enumerate(in_targets)}
        tout = {t: i for i, t in enumerate(out_targets)}
        n_in, n_out, w_off, ws = [], [], [], []
        ent_in, ent_out = [], []
        woff = 0
        for s in valid:
            irreps = tuple(irreps_per_leg[k][grid[s, k]] for k in range(ndim))
            secs_in, secs_out, U = cls._sigma_unitary(irreps, signature, nrr, axes, nrr_out)
            for secs, tmap, ent in ((secs_in, tin, ent_in), (secs_out, tout, ent_out)):
                for J, nr, nc in secs:
                    a = np.repeat(np.arange(nr, dtype=np.int64), nc)
                    b = np.tile(np.arange(nc, dtype=np.int64), nr)
                    ent.append(np.stack([np.full_like(a, tmap[J]), a, b]))
            pad = int(_pad8(U.shape[0]))
            wt = np.zeros((U.shape[1], pad))
            wt[:, : U.shape[0]] = U.T
            ws.append(wt.reshape(-1))
            n_in.append(U.shape[1])
            n_out.append(U.shape[0])
            w_off.append(woff)
            woff += wt.size
        st = _PermuteStructure()
        st.in_targets = np.array(in_targets, dtype=np.int64)
        st.out_targets = np.array(out_targets, dtype=np.int64)
        st.sig_idx = grid[valid]
        st.ir_in, st.ic_in = ir_in[valid], ic_in[valid]
        st.ir_out, st.ic_out = ir_out[valid], ic_out[valid]
        st.idr_in, st.idc_in, st.idr_out, st.idc_out = idr_in, idc_in, idr_out, idc_out
        st.n_in = np.array(n_in, dtype=np.int64)
        st.n_out = np.array(n_out, dtype=np.int64)
        st.n_out_pad = _pad8(st.n_out)
        st.w_off = np.array(w_off, dtype=np.int64)
        st.W = np.concatenate(ws) if ws else np.zeros((0,))
        ei = np.concatenate(ent_in, axis=1) if ent_in else np.zeros((3, 0), dtype=np.int64)
        eo = np.concatenate(ent_out, axis=1) if ent_out else np.zeros((3, 0), dtype=np.int64)
        st.in_t, st.in_a, st.in_b = ei
        st.out_t, st.out_a, st.out_b = eo
        st.bind(ndim, nrr, nrr_out, axes, nper)
        cls._permute_structures[key] = st
        plan_store.note_built(time.perf_counter() - t_build)
        plan_store.store(skey, st.to_arrays())
        return st

    def _block_strides(self):
        """Row/column strides of each stored block (same memory order for all)."""
        if self._nblocks == 0:
            return np.zeros((0,), np.int64), np.zeros((0,), np.int64)
        if self._fortran:
            return np.ones((self._nblocks,), np.int64), self._bshapes[:, 0].copy()
        return self._bshapes[:, 1].copy(), np.ones((self._nblocks,), np.int64)

    def _present_map(self, targets):
        """target index -> stored block index (-1 when the block is missing)."""
        pos = np.full((len(targets),), -1, dtype=np.int64)
        if self._nblocks:
            ti = np.searchsorted(targets, self._block_irreps)
            assert not _debug_config["debug_checks"] or (
                (ti < len(targets)).all() and (targets[ti] == self._block_irreps).all()
            )
            pos[ti] = np.arange(self._nblocks)
        return pos

    def _permute_plan(self, axes, nrr_out):
        cls = type(self)
        reps = self._row_reps + self._col_reps
        nrr = self._nrr
        axes = tuple(int(a) for a in axes)
        # the key hashes raw bytes (representations: degeneracies and irreps of
        # every leg); the irreps tuples are only built for a new plan
        key = (
            cls, tuple(np.ascontiguousarray(r, dtype=np.int64).tobytes() for r in reps),
            self._signature.tobytes(), nrr, axes, nrr_out,
            self._block_irreps.tobytes(), self._boffs.tobytes(), self._bshapes.tobytes(),
            self._fortran,
        )
        with cls._plans_lock:
            hit = cls._plans.get(key)
            if hit is not None:
                cls._plans.move_to_end(key)
                return hit
        # irreps-only structure, found by the irreps' bytes; the tuples the
        # structure is built from are formed only the first time
        skey = (cls, tuple(np.ascontiguousarray(r[1], dtype=np.int64).tobytes() for r in reps),
                key[2], nrr, axes, nrr_out)
        st = cls._structures_by_bytes.get(skey)
        if st is None:
            irreps_per_leg = tuple(tuple(int(x) for x in r[1]) for r in reps)
            sig = tuple(bool(s) for s in self._signature)
            st = cls._structures_by_bytes[skey] = cls._permute_structure(irreps_per_leg, sig, nrr, axes, nrr_out)
        degen = np.concatenate([np.asarray(r[0], dtype=np.int64) for r in reps])
        rs_b, cs_b = self._block_strides()
        plan, shapes = _ops.lie_plan_build(
            st.pattern, degen, self._present_map(st.in_targets), self._boffs, rs_b, cs_b, st.weights,
            self._data.numel(),
        )
        hit = (plan, shapes, st.out_targets)
        with cls._plans_lock:
            cls._plans[key] = hit
            if len(cls._plans) > cls._PLAN_CACHE_SIZE:
                cls._plans.popitem(last=False)
        return hit

    def _permute_data(self, axes, nrr):
        plan, shapes, irreps = self._permute_plan(axes, nrr)
        out = _ops.empty(plan.dst_size)
        if plan.dst_size:
            plan.execute(self._data, out)
        return out, shapes, irreps.copy()

    def transpose(self):
        """Matrix transpose: a permute of rows and columns (lie:1111-1114); the
        structural signs make it a data movement, not a relabelling."""
        axes = tuple(range(self._nrr, self._ndim)) + tuple(range(self._nrr))
        return self._permuted(axes, self._ndim - self._nrr)

    # dense <-> blocks -------------------------------------------------------------
    @classmethod
    def _dense_structure(cls, irreps_per_leg, signature, nrr, targets):
        key = (cls, irreps_per_leg, signature, nrr, targets)
        st = cls._dense_structures.get(key)
        if st is not None:
            return st
        ndim = len(irreps_per_leg)
        nper = np.array([len(ir) for ir in irreps_per_leg], dtype=np.int64)
        grid = _grid(nper)
        f_r = tuple(bool(s) for s in signature[:nrr])
        f_c = tuple(not s for s in signature[nrr:])
        idr = cls._side_internal(irreps_per_leg[:nrr], f_r, targets)
        idc = cls._side_internal(irreps_per_leg[nrr:], f_c, targets)
        ir = np.ravel_multi_index(tuple(grid[:, k] for k in range(nrr)), tuple(nper[:nrr])) if nrr else 0
        ic = np.ravel_multi_index(tuple(grid[:, k] for k in range(nrr, ndim)), tuple(nper[nrr:]))
        valid = np.flatnonzero((idr[ir] * idc[ic]).sum(axis=1) > 0)
        tpos = {t: i for i, t in enumerate(targets)}
        st = _DenseStructure()
        st.targets = np.array(targets, dtype=np.int64)
        st.sig_idx = grid[valid]
        st.ir, st.ic = ir[valid], ic[valid]
        st.idr, st.idc = idr, idc
        st.m_idx, st.proj, st.sec_t, st.sec_a, st.sec_b, st.n_rows = [], [], [], [], [], []
        for s in valid:
            irreps = tuple(irreps_per_leg[k][grid[s, k]] for k in range(ndim))
            secs, P = cls._sigma_projector(irreps, signature, nrr, targets=set(targets))
            live = np.flatnonzero(np.any(P != 0.0, axis=1))
            dims = tuple(cls.irrep_dimension(d) for d in irreps)
            st.m_idx.append(np.stack(np.unravel_index(live, dims), axis=1).astype(np.int64))
            st.proj.append(P[live])
            t, a, b = [], [], []
            for J, nr, nc in secs:
                t.append(np.full((nr * nc,), tpos[J], dtype=np.int64))
                a.append(np.repeat(np.arange(nr, dtype=np.int64), nc))
                b.append(np.tile(np.arange(nc, dtype=np.int64), nr))
            st.sec_t.append(np.concatenate(t))
            st.sec_a.append(np.concatenate(a))
            st.sec_b.append(np.concatenate(b))
        cls._dense_structures[key] = st
        return st

    @classmethod
    def _dense_plan(cls, reps, signature, nrr, targets, block_offs, block_rs, block_cs, to_dense):
        """Plan between the dense matrix (row-major, rows = row legs) and the
        blocks of `targets` (at block_offs with strides block_rs/block_cs)."""
        ndim = len(reps)
        irreps_per_leg = tuple(tuple(int(x) for x in r[1]) for r in reps)
        sig = tuple(bool(s) for s in signature)
        st = cls._dense_structure(irreps_per_leg, sig, nrr, tuple(int(t) for t in targets))
        degens = [np.asarray(r[0], dtype=np.int64) for r in reps]
        dims_leg = [np.array([cls.irrep_dimension(x) for x in r[1]], dtype=np.int64) for r in reps]
        shifts = [_exclusive_cumsum0(d * g) for d, g in zip(dims_leg, degens, strict=True)]
        tot = np.array([int((d * g).sum()) for d, g in zip(dims_leg, degens, strict=True)], dtype=np.int64)
        ncols_dense = int(np.prod(tot[nrr:]))
        rstride = np.ones((ndim,), dtype=np.int64)
        for k in range(nrr - 2, -1, -1):
            rstride[k] = rstride[k + 1] * tot[k + 1]
        for k in range(ndim - 2, nrr - 1, -1):
            rstride[k] = rstride[k + 1] * tot[k + 1]
        edr = _outer(degens[:nrr])
        edc = _outer(degens[nrr:])
        rstart = _exclusive_cumsum0(edr[:, None] * st.idr)
        cstart = _exclusive_cumsum0(edc[:, None] * st.idc)
        tasks, rows, digs, ws = [], [], [], []
        nrow = woff = 0
        for s in range(st.sig_idx.shape[0]):
            idx = st.sig_idx[s]
            ir, ic = st.ir[s], st.ic[s]
            m = st.m_idx[s]
            shift = np.array([shifts[k][idx[k]] for k in range(ndim)], dtype=np.int64)
            lin = (shift[None, :] + m) * rstride[None, :]
            d_base = lin[:, :nrr].sum(axis=1) * ncols_dense + lin[:, nrr:].sum(axis=1)
            dense_rows = np.stack([d_base, np.full_like(d_base, ncols_dense), np.ones_like(d_base)], axis=1)
            t, a, b = st.sec_t[s], st.sec_a[s], st.sec_b[s]
            brow = rstart[ir, t] + a * edr[ir]
            bcol = cstart[ic, t] + b * edc[ic]
            blk_rows = np.stack([block_offs[t] + brow * block_rs[t] + bcol * block_cs[t], block_rs[t], block_cs[t]], axis=1)
            deg = np.array([degens[k][idx[k]] for k in range(ndim)], dtype=np.int64)
            dim = np.array([dims_leg[k][idx[k]] for k in range(ndim)], dtype=np.int64)
            d = np.zeros((ndim, 5), dtype=np.int64)
            d[:, 0] = deg
            dense_r = np.where(np.arange(ndim) < nrr, dim * rstride, 0)
            dense_c = np.where(np.arange(ndim) >= nrr, dim * rstride, 0)
            blk_r = np.zeros((ndim,), dtype=np.int64)
            blk_c = np.zeros((ndim,), dtype=np.int64)
            blk_r[:nrr] = _suffix_products(deg[None, :nrr])[0]
            blk_c[nrr:] = _suffix_products(deg[None, nrr:])[0]
            P = st.proj[s]
            if to_dense:
                in_rows, out_rows = blk_rows, dense_rows
                d[:, 1], d[:, 2], d[:, 3], d[:, 4] = blk_r, blk_c, dense_r, dense_c
                W = P  # (n_dense, S)
            else:
                in_rows, out_rows = dense_rows, blk_rows
                d[:, 1], d[:, 2], d[:, 3], d[:, 4] = dense_r, dense_c, blk_r, blk_c
                W = P.T / np.array([cls.irrep_dimension(j) for j in st.targets[t]])[:, None]
            n_out, n_in = W.shape
            pad = int(_pad8(n_out))
            wt = np.zeros((n_in, pad))
            wt[:, :n_out] = W.T
            tasks.append([int(deg.prod()), woff, n_in, n_out, pad, nrow, ndim, s * ndim])
            rows.append(in_rows)
            rows.append(out_rows)
            digs.append(d)
            ws.append(wt.reshape(-1))
            nrow += n_in + n_out
            woff += wt.size
        if not tasks:
            return None
        return (
            np.array(tasks, dtype=np.int64),
            np.concatenate(rows),
            np.concatenate(digs),
            np.concatenate(ws),
        )

    @classmethod
    def _blocks_from_dense(cls, arr, row_reps, col_reps, signature):
        """Dense -> blocks: per elementary block, CG projection / irrep dimension
        (lie:856-967). The dense input is uploaded once; projection runs on the
        device."""
        reps = tuple(row_reps) + tuple(col_reps)
        nrr = len(row_reps)
        block_irreps, shapes = cls.get_block_sizes(row_reps, col_reps, signature)
        sizes = shapes[:, 0] * shapes[:, 1]
        offs = _exclusive_cumsum0(sizes)
        out = _ops.zeros(int(sizes.sum()))
        tables = cls._dense_plan(
            reps, signature, nrr, block_irreps, offs, shapes[:, 1].copy(), np.ones_like(offs), to_dense=False
        )
        if tables is not None:
            src = _dense_source(arr)
            plan = _ops.LiePlan(*tables, src.numel(), out.numel())
            plan.execute(src, out)
        return out, shapes, block_irreps

    def _tomatrix(self):
        """Blocks -> dense (lie:969-1109), zero outside the invariant subspace."""
        return _ops.download(self._tomatrix_device()).reshape(self.matrix_shape)

    def _tomatrix_device(self):
        """Dense row-major matrix as a flat device buffer."""
        reps = self._row_reps + self._col_reps
        n = int(np.prod(self._shape))
        rs_b, cs_b = self._block_strides()
        tables = type(self)._dense_plan(
            reps, self._signature, self._nrr, self._block_irreps, self._boffs, rs_b, cs_b, to_dense=True
        )
        dense = _ops.zeros(n)
        if tables is not None:
            plan = _ops.LiePlan(*tables, self._data.numel(), n)
            plan.execute(self._data, dense)
        return dense

    def update_signature(self, sign_update):
        raise NotImplementedError("signature update of Lie-group tensors (lie:1136-1176)")
