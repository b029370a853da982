"""Reduced density matrices from a converged environment (ctmrg/rdm.py:1-189).

The contraction orders (and leg conventions of the returned rdm: bra legs as
rows, ket legs as columns) follow the reference so that energies computed as
Tr(rho h) are interchangeable.
"""


def rdm_1x1(C1, T1, C2, T4, A, T2, C4, T3, C3):
    """Single-site rdm (rdm.py:1-27)."""
    right = C3.transpose() @ T2.permute((1,), (0, 2, 3))
    right = right.permute((0, 2, 3), (1,)) @ C2
    down = C4 @ T3.permute((3,), (0, 1, 2))
    rdm = T1.permute((0, 1, 2), (3,)) @ C1
    rdm = rdm @ T4.permute((0,), (1, 2, 3))
    rdm = rdm.permute((0, 2, 4, 5), (1, 3)) @ A.permute((2, 5), (0, 1, 3, 4))
    down = down.permute((0, 1, 2), (3,)) @ right.permute((0,), (1, 2, 3))
    down = down.permute((2, 4), (5, 3, 0, 1)) @ rdm.permute((0, 6, 3, 7), (1, 2, 4, 5))
    down = down.permute((4,), (2, 3, 1, 0, 5))
    rdm = down @ A.permute((0,), (2, 5, 3, 4, 1)).dagger()
    rdm /= rdm.trace()
    return rdm


def contract_open_corner(C1, T1, T4, A):
    """Upper-left corner with open physical legs (rdm.py:30-80)."""
    ul = T1.permute((1, 2, 0), (3,)) @ C1
    ul = ul @ T4.permute((0,), (1, 2, 3))
    ul = ul.permute((0, 3), (1, 4, 2, 5))
    ul = A.permute((1, 0, 3, 4), (2, 5)) @ ul
    ul = ul.permute((0, 4, 5), (1, 2, 3, 6, 7))
    ul = A.permute((1, 2, 5), (0, 3, 4)).dagger() @ ul
    return ul.permute((3, 0, 4, 1, 6), (5, 2, 7))


def contract_open_corner_mirror(T1, C2, A, T2):
    """Upper-right mirror of contract_open_corner (rdm.py:83-100)."""
    ur = C2 @ T1.permute((0,), (1, 2, 3))
    ur = ur.transpose() @ T2.permute((0,), (2, 3, 1))
    ur = ur.permute((0, 3), (1, 4, 2, 5))
    ur = A.permute((1, 0, 4, 5), (2, 3)) @ ur
    ur = ur.permute((0, 4, 5), (1, 2, 3, 6, 7))
    return A.permute((1, 2, 3), (0, 4, 5)).dagger() @ ur


def rdm_1x2(C1, T1l, T1r, C2, T4, Al, Ar, T2, C4, T3l, T3r, C3):
    """Two horizontally adjacent sites (rdm.py:103-136)."""
    left = _half_1x2(C1, T1l, T4, Al, C4, T3l)
    right = T2.permute((0, 2, 3), (1,)) @ C3
    right = right.permute((0, 3), (1, 2))
    right = contract_open_corner_mirror(T1r, C2, Ar, right)
    right = right.permute((3, 0, 5, 2, 6), (4, 1, 7))
    right = right @ T3r.permute((0, 1, 2), (3,))
    right = right.permute((2, 3, 4, 5), (0, 1))
    rdm = (left @ right).permute((1, 3), (0, 2))
    rdm /= rdm.trace()
    return rdm


def rdm_2x1(C1, T1, C2, T4u, Au, T2u, T4d, Ad, T2d, C4, T3, C3):
    """Two vertically adjacent sites through rdm_1x2 on rotated legs (rdm.py:171-189)."""
    return rdm_1x2(
        C2,
        T2u.permute((1, 2, 3), (0,)),
        T2d.permute((1,), (2, 3, 0)),
        C3.transpose(),
        T1,
        Au.permute((0, 1), (3, 4, 5, 2)),
        Ad.permute((0, 1), (3, 4, 5, 2)),
        T3.permute((2, 3, 0), (1,)),
        C1,
        T4u.permute((1, 2, 3), (0,)),
        T4d.permute((1, 2, 3), (0,)),
        C4.transpose(),
    )


def _half_1x2(C1, T1l, T4, Al, C4, T3l):
    """Left half of a horizontal pair with both physical legs open: legs
    (bra, ket) x (T1 chi, C? chi, ...) ready to meet the right half."""
    left = contract_open_corner(C1, T1l, T4.permute((0, 1, 2), (3,)) @ C4, Al)
    left = left @ T3l.permute((0, 1, 3), (2,))
    return left.permute((0, 1), (2, 3, 4, 5))


def rdm_symmetrized_1x2(C1, T1l, T4, Al, C4, T3l):
    """Horizontal pair in an environment symmetric under the reflection that
    maps the left site onto the right one (rdm.py:139-153): the right half is
    the mirrored left half."""
    left = _half_1x2(C1, T1l, T4, Al, C4, T3l)
    rdm = (left @ left.permute((3, 2, 5, 4), (1, 0))).permute((1, 3), (0, 2))
    rdm /= rdm.trace()
    return rdm


def rdm_symmetrized_2x1(C1, T1, C2, T4u, Au, T2u):
    """Vertical pair through rdm_symmetrized_1x2 on rotated legs (rdm.py:156-168)."""
    return rdm_symmetrized_1x2(
        C2,
        T2u.permute((1, 2, 3), (0,)),
        T1,
        Au.permute((0, 1), (3, 4, 5, 2)),
        C1,
        T4u.permute((1, 2, 3), (0,)),
    )


def rdm_diag_dr(C1, T1l, ur, T4u, Aul, dl, Adr, T2d, T3r, C3):
    """Next-nearest neighbours on the down-right diagonal (rdm.py:192-245):
    open upper-left corner, closed enlarged ur and dl corners, open down-right
    corner built as the mirror of the upper-right one."""
    rdm = contract_open_corner(C1, T1l, T4u, Aul) @ dl
    rdm = ur @ rdm.permute((2, 3, 4), (0, 1, 5, 6, 7))
    rdm = rdm.permute((3, 4), (0, 1, 2, 5, 6, 7))
    dr = contract_open_corner_mirror(
        T2d.permute((1, 2, 3), (0,)),
        C3.transpose(),
        Adr.permute((0, 1, 3, 4), (5, 2)),
        T3r.permute((2, 3), (0, 1)),
    )
    rdm = rdm @ dr.permute((5, 2, 6, 4, 1, 7), (3, 0))
    rdm = rdm.permute((1, 3), (0, 2))
    rdm /= rdm.trace()
    return rdm


def rdm_diag_ur(ul, T1r, C2, Aur, T2u, T4d, Adl, dr_T, C4, T3l):
    """Next-nearest neighbours on the up-right diagonal through rdm_diag_dr on
    the lattice rotated by a quarter turn (rdm.py:248-272); dr_T is the
    enlarged down-right corner in its transposed storage order."""
    rdm = rdm_diag_dr(
        C4,
        T4d,
        ul,
        T3l.permute((3,), (0, 1, 2)),
        Adl.permute((0, 1), (5, 2, 3, 4)),
        dr_T,
        Aur.permute((0, 1), (5, 2, 3, 4)),
        T1r.permute((3,), (0, 1, 2)),
        T2u.permute((2, 3, 0), (1,)),
        C2.transpose(),
    )
    return rdm.permute((1, 0), (3, 2))
