"""Enlarged-corner and norm contractions of CTMRG.

The leg conventions (and therefore the sequence of permute / @ on every
contraction) are part of the CTMRG interface: environment tensors must carry the
same legs in the same order as frostspin's (ctmrg/ctm_contract.py:7-162) so that
environments, projectors and reduced density matrices are interchangeable.
Each step is one device permute (planned gather) or one grouped MFMA GEMM.

Site tensor A legs: (physical, ancilla, up, right, down, left); C/T legs follow
the clockwise convention documented in abstract_ctmrg.py:26-42.
"""

import weakref

# Permuted copies of site tensors, kept alongside the tensor they come from:
# the site tensors of an environment do not change between set_tensors calls
# and every move needs the same few permutations of them (two per enlarged
# corner or renormalized edge). forget_site_permutes() drops them all.
_SITE_PERMUTES = weakref.WeakKeyDictionary()


def site_permute(A, row_axes, col_axes):
    """A.permute(row_axes, col_axes) for a site tensor (or a permuted copy of
    one), computed once and kept while A lives."""
    kept = _SITE_PERMUTES.get(A)
    if kept is None:
        kept = _SITE_PERMUTES[A] = {}
    key = (tuple(row_axes), tuple(col_axes))
    t = kept.get(key)
    if t is None:
        t = kept[key] = A.permute(row_axes, col_axes)
    return t


def forget_site_permutes():
    _SITE_PERMUTES.clear()


def contract_enlarged_corner(C1, T1, T4, A):
    """Upper-left template C1-T1 // T4-A-A*; other corners are fed rotated legs
    (ctm_contract.py:7-61). Memory peak 2*a*d*chi^2*D^4 coefficients (dense
    equivalent) at the fourth step."""
    ul = T1 @ C1
    ul = ul @ T4
    ul = A @ ul.permute((0, 3), (1, 4, 2, 5))
    ul = ul.permute((0, 1, 4, 5), (2, 3, 6, 7))
    ul = site_permute(A, (0, 1, 4, 5), (2, 3)).dagger() @ ul
    return ul.permute((2, 0, 4), (3, 1, 5))


def contract_ul(C1, T1, T4, A):
    return contract_enlarged_corner(
        C1, T1.permute((1, 2, 0), (3,)), T4.permute((0,), (1, 2, 3)), site_permute(A, (0, 1, 3, 4), (2, 5))
    )


def contract_ur(T1, C2, A, T2):
    return contract_enlarged_corner(
        C2, T2.permute((2, 3, 1), (0,)), T1.permute((0,), (1, 2, 3)), site_permute(A, (0, 1, 4, 5), (3, 2))
    )


def contract_dr(A, T2, T3, C3):
    """Transposed with respect to clockwise order (dr is contracted as dl-10-dr)."""
    return contract_enlarged_corner(
        C3.transpose(), T3.permute((0, 1, 3), (2,)), T2.permute((1,), (2, 3, 0)),
        site_permute(A, (0, 1, 5, 2), (4, 3)),
    )


def contract_dl(T4, A, C4, T3):
    return contract_enlarged_corner(
        C4, T4.permute((1, 2, 0), (3,)), T3.permute((3,), (0, 1, 2)), site_permute(A, (0, 1, 2, 3), (5, 4))
    )


def contract_C1234(C1, C2, C4, C3):
    """Four corners closed on themselves (ctm_contract.py:103-106)."""
    return (C2 @ C1).full_contract(C4 @ C3.transpose())


def contract_T1T3(C1, T1, C2, C4, T3, C3):
    """Corners plus the up/down edges (ctm_contract.py:109-127)."""
    left = T1.permute((0, 1, 2), (3,)) @ (C1 @ C4)
    right = T3.permute((3, 0, 1), (2,)) @ (C3.transpose() @ C2)
    return left.full_contract(right.permute((0,), (3, 1, 2)))


def contract_T2T4(C1, C2, T4, T2, C4, C3):
    """Corners plus the left/right edges (ctm_contract.py:130-138)."""
    left = (C2 @ C1) @ T4.permute((0,), (1, 2, 3))
    left = left.permute((0, 1, 2), (3,)) @ (C4 @ C3.transpose())
    return left.full_contract(T2.permute((1,), (0, 2, 3)))


def contract_norm(C1, T1, C2, T4, A, T2, C4, T3, C3):
    """Full 3x3 network around one site (ctm_contract.py:141-162)."""
    ul = contract_ul(C1, T1, T4, A)
    right = (C2.transpose() @ T2.permute((0,), (2, 3, 1))).permute((3,), (1, 2, 0))
    up = right @ ul
    down = (C3 @ T3.permute((2,), (0, 1, 3))).permute((3,), (1, 2, 0))
    down = (C4 @ down).permute((1, 2, 0), (3,))
    return up.full_contract(down)
