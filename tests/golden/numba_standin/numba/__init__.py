"""Pure-Python stand-in for the four numba symbols frostspin uses.

Used only by tests/golden/make_reference_fixtures.py, in the build container,
to import the reference checkout (which needs numba at import time) and write
small golden vectors. It never travels to a GPU run and the product never
imports it.

Symbols provided (the complete set the reference uses):
  numba.njit / numba.njit(parallel=True) -> identity decorator
  numba.prange                            -> range
  numba.typed.List                        -> list
  numba.np.unsafe.ndarray.to_fixed_tuple  -> tuple(a[:n])
"""

from types import SimpleNamespace


def njit(*args, **kwargs):
    if len(args) == 1 and callable(args[0]) and not kwargs:
        return args[0]

    def deco(fn):
        return fn

    return deco


prange = range


def _typed_list(*args):
    return list(*args)


def _to_fixed_tuple(a, n):
    return tuple(int(x) for x in a[:n])


typed = SimpleNamespace(List=_typed_list)
np = SimpleNamespace(unsafe=SimpleNamespace(ndarray=SimpleNamespace(to_fixed_tuple=_to_fixed_tuple)))
