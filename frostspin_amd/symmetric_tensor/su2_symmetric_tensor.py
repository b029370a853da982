"""SU(2) symmetric tensors (su2_symmetric_tensor.py:130-338).

Irreps are labelled by their dimension 2s+1; a representation is an int array
(2, n) of (degeneracies, irreps). Fusion trees are left-to-right chains of
Clebsch-Gordan tensors (su2:213-252); a leg with a flipped signature carries
the spin-reversal map (su2:87-91).
"""

import numpy as np

from .. import _ops
from ..groups import su2_clebsch_gordan as cg
from .lie_group_symmetric_tensor import LieGroupSymmetricTensor
from .tools import check_norm, symmetric_tensor_types


def _fuse_pair(rep1, rep2):
    """(degen, irreps) of rep1 x rep2 (su2:109-119)."""
    degen = np.zeros((int(rep1[1][-1] + rep2[1][-1] - 1),), dtype=np.int64)
    for d1, i1 in zip(rep1[0], rep1[1], strict=True):
        for d2, i2 in zip(rep2[0], rep2[1], strict=True):
            degen[abs(int(i1) - int(i2)) : int(i1 + i2) - 1 : 2] += int(d1) * int(d2)
    nz = degen.nonzero()[0]
    return np.array([degen[nz], nz + 1], dtype=np.int64)


class SU2SymmetricTensor(LieGroupSymmetricTensor):
    _symmetry = "SU2"
    _trees = {}

    @staticmethod
    def singlet():
        return np.ones((2, 1), dtype=np.int64)

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

    @staticmethod
    def representation_dimension(rep):
        return int(np.asarray(rep[0]) @ np.asarray(rep[1]))

    @staticmethod
    def irrep_dimension(irr):
        return int(irr)

    @classmethod
    def clebsch_gordan_dic(cls):
        """Computed-on-demand CG tensors keyed (d1, d2, d3)."""
        return _CGView()

    @classmethod
    def reload_clebsch_gordan(cls, savefile, *, verbosity=0):
        """Use coefficients from a table file (su2:321-335 format) instead of
        the computed ones; clears every structural cache."""
        cg.load_su2_cg(savefile, verbosity=verbosity)
        for cache in (cls._trees, cls._unitary_cache, cls._permute_structures,
                      cls._dense_structures, cls._side_cache, cls._plans):
            cache.clear()

    @classmethod
    def _elementary_trees(cls, irreps, flags):
        """Left-to-right CG chain of one elementary block (compute_fusion_tensor
        and compute_clebsch_gordan_tree, su2:14-106, 213-252): the internal index
        of an output irrep enumerates the fusion paths, latest intermediate
        irrep most significant, each intermediate ascending."""
        key = (tuple(int(d) for d in irreps), tuple(bool(f) for f in flags))
        hit = cls._trees.get(key)
        if hit is not None:
            return hit
        dims, flags = key
        zs = [cg.spin_reversal(d) if f else None for d, f in zip(dims, flags, strict=True)]
        if len(dims) == 1:
            d = dims[0]
            t = zs[0] if zs[0] is not None else np.eye(d)
            out = {d: t.reshape(d, 1, d)}
        else:
            d0, d1 = dims[0], dims[1]
            cur = {}
            for c in range(abs(d0 - d1) + 1, d0 + d1, 2):
                p = cg.clebsch_gordan(d0, d1, c)
                if zs[0] is not None:
                    p = np.tensordot(zs[0], p, axes=(1, 0))
                if zs[1] is not None:
                    p = np.tensordot(zs[1], p, axes=(1, 1)).transpose(1, 0, 2)
                cur[c] = p.reshape(d0 * d1, 1, c)
            for k in range(2, len(dims)):
                b = dims[k]
                nxt = {}
                for c in sorted(cur):
                    t = cur[c]  # (N, n_c, c)
                    for e in range(abs(c - b) + 1, c + b, 2):
                        p = cg.clebsch_gordan(c, b, e)
                        if zs[k] is not None:
                            p = np.tensordot(zs[k], p, axes=(1, 1)).transpose(1, 0, 2)
                        piece = np.einsum("Nic,cbe->Nbie", t, p, optimize=True)
                        nxt.setdefault(e, []).append(piece.reshape(t.shape[0] * b, t.shape[1], e))
                cur = {e: np.concatenate(v, axis=1) for e, v in nxt.items()}
            out = dict(sorted(cur.items()))
        for v in out.values():
            v.flags.writeable = False
        cls._trees[key] = out
        return out

    def toSU2(self):
        return self

    def toabelian(self):
        return self.toU1()

    def toU1(self):
        """Cast through the dense array with 2*Sz charges, Sz descending inside
        each multiplet (su2:257-278)."""
        from .u1_symmetric_tensor import U1SymmetricTensor

        reps = []
        for r in self._row_reps + self._col_reps:
            sz = []
            for d, irr in zip(r[0], r[1], strict=True):
                sz.append(np.tile(np.arange(irr - 1, -irr - 1, -2, dtype=np.int8), int(d)))
            reps.append(np.concatenate(sz).astype(np.int8))
        out = U1SymmetricTensor.from_array(
            self.toarray(), reps[: self._nrr], reps[self._nrr :], self._signature
        )
        assert check_norm(self, out)
        return out


    def toO2(self):
        """Cast to O(2) (su2:280-318): every leg's basis is reordered from SU(2)
        multiplets to O(2) sectors (0odd, 0even, then +n / -n blocks with n =
        2|Sz|) and gets the O(2) reflection signs; one signed row/column
        permutation of the dense matrix on the device, then O(2) from_array."""
        from .o2_symmetric_tensor import O2SymmetricTensor

        legs = [_o2_leg_cast(tuple(r[0].tolist()), tuple(r[1].tolist())) for r in self._row_reps + self._col_reps]
        nrr = self._nrr
        rows, cols = self.matrix_shape

        def composed(parts):
            shape = [len(p[0]) for p in parts]
            grid = np.indices(shape).reshape(len(shape), -1) if shape else np.zeros((0, 1), np.int64)
            src = np.zeros(grid.shape[1], dtype=np.int64)
            sign = np.ones(grid.shape[1])
            for k, (perm, sg, _) in enumerate(parts):
                src = src * shape[k] + perm[grid[k]]
                sign *= sg[perm[grid[k]]]
            return src, sign

        rsrc, rsign = composed(legs[:nrr])
        csrc, csign = composed(legs[nrr:])
        idx = np.zeros((2 * rows + 2 * cols,), dtype=np.int64)
        w = np.zeros((2 * rows + 2 * cols,))
        idx[0 : 2 * rows : 2] = rsrc
        w[0 : 2 * rows : 2] = rsign
        idx[2 * rows :: 2] = csrc
        w[2 * rows :: 2] = csign
        dense = self._tomatrix_device()
        out = _ops.empty(rows * cols)
        _ops.gather2d(
            [(rows, cols, rows, cols, 0, cols, 1, 0, cols, 1, 0, 0, 2 * rows, 2 * rows, 0.0)], idx, w, dense, out
        )
        o2_reps = [leg[2] for leg in legs]
        res = O2SymmetricTensor.from_array(out.reshape(self._shape), o2_reps[:nrr], o2_reps[nrr:], self._signature)
        assert check_norm(self, res)
        return res


def _o2_leg_cast(degen, irreps):
    """(source index, sign, O(2) representation) of one leg (su2:295-318)."""
    sz, sz_o2, flip = [], [], []
    for d, irr in zip(degen, irreps, strict=True):
        s = np.arange(irr - 1, -irr - 1, -2, dtype=np.int64)
        so2 = np.abs(s)
        f = np.zeros(s.shape, dtype=bool)
        if irr % 2:
            f[(s % 4 == (irr + 1) % 4) & (s < 0)] = True
            if irr % 4 == 3:  # odd integer spin: its Sz = 0 state is 0odd
                so2[irr // 2] = -1
        else:
            f[(s % 4 == (irr - 1) % 4) & (s < 0)] = True
        for _ in range(d):
            sz.append(s)
            sz_o2.append(so2)
            flip.append(f)
    sz = np.concatenate(sz)
    sz_o2 = np.concatenate(sz_o2)
    sign = 1.0 - 2.0 * np.concatenate(flip)
    swap = np.argsort(-sz, kind="stable")
    swap = swap[np.argsort(sz_o2[swap], kind="stable")]
    o2_irreps, o2_degen = np.unique(sz_o2, return_counts=True)
    o2_degen[o2_irreps > 0] //= 2
    return swap, sign, np.array([o2_degen, o2_irreps], dtype=np.int64)


class _CGView:
    """Mapping-like access to CG tensors: view[(d1, d2, d3)]."""

    def __getitem__(self, key):
        try:
            return cg.clebsch_gordan(*key)
        except ValueError:
            raise KeyError(key) from None


symmetric_tensor_types["SU2"] = SU2SymmetricTensor
