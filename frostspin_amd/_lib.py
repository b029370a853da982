"""ctypes binding of libfrostspin_amd.so (the C-ABI declared in include/frostspin_amd.h).

The library is loaded at import. A missing library is a hard error: there is no
host fallback for any device operation.
"""

import ctypes
import os
import sys
from pathlib import Path

import numpy as np

_LIB_PATH = Path(__file__).resolve().parent / "libfrostspin_amd.so"

FS_NORM_PARTIALS = 1024
FS_PERMUTE_MAX_DIGITS = 12
FS_PERMUTE_TILE_RECORD = 5 + 5 * FS_PERMUTE_MAX_DIGITS

c_i64 = ctypes.c_int64
c_int = ctypes.c_int
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p
c_i64p = ctypes.POINTER(ctypes.c_int64)
c_i32p = ctypes.POINTER(ctypes.c_int32)
c_i8p = ctypes.POINTER(ctypes.c_int8)
c_u8p = ctypes.POINTER(ctypes.c_uint8)
c_dblp = ctypes.POINTER(ctypes.c_double)


class GemmGroup(ctypes.Structure):
    _fields_ = [
        ("m", c_i64),
        ("n", c_i64),
        ("k", c_i64),
        ("a_off", c_i64),
        ("a_rs", c_i64),
        ("a_cs", c_i64),
        ("b_off", c_i64),
        ("b_rs", c_i64),
        ("b_cs", c_i64),
        ("c_off", c_i64),
        ("c_ld", c_i64),
    ]


class LiePattern(ctypes.Structure):
    """fs_lie_pattern: irreps-only data of one permute pattern (int64 arrays)."""

    _fields_ = [(n, c_i64) for n in ("ndim", "nrr", "nrr_out", "ntasks", "nt_in", "nt_out")] + [
        (n, c_i64p)
        for n in (
            "axes", "nirr", "sig_idx", "ir_in", "ic_in", "ir_out", "ic_out", "n_in", "n_out",
            "w_off", "n_out_pad", "idr_in", "idc_in", "idr_out", "idc_out",
            "in_t", "in_a", "in_b", "out_t", "out_a", "out_b",
        )
    ]


class SvdDesc(ctypes.Structure):
    _fields_ = [
        ("m", c_i64),
        ("n", c_i64),
        ("a_off", c_i64),
        ("a_rs", c_i64),
        ("a_cs", c_i64),
        ("u_off", c_i64),
        ("s_off", c_i64),
        ("vt_off", c_i64),
        ("work_off", c_i64),
    ]


class QrDesc(ctypes.Structure):
    _fields_ = [
        ("m", c_i64),
        ("n", c_i64),
        ("a_off", c_i64),
        ("a_rs", c_i64),
        ("a_cs", c_i64),
        ("q_off", c_i64),
        ("r_off", c_i64),
        ("work_off", c_i64),
    ]


class CholDesc(ctypes.Structure):
    _fields_ = [("l", c_i64), ("g_off", c_i64), ("r_off", c_i64), ("shift_rel", c_dbl)]


class CholInvDesc(ctypes.Structure):
    _fields_ = [
        ("l", c_i64),
        ("g_off", c_i64),
        ("g_stride", c_i64),
        ("n_part", c_i64),
        ("r_off", c_i64),
        ("rinv_off", c_i64),
        ("shift_rel", c_dbl),
    ]


class TriInvDesc(ctypes.Structure):
    _fields_ = [("l", c_i64), ("r_off", c_i64), ("x_off", c_i64)]


class CholQrDesc(ctypes.Structure):
    _fields_ = [
        ("m", c_i64),
        ("l", c_i64),
        ("y_off", c_i64),
        ("y_rs", c_i64),
        ("y_cs", c_i64),
        ("q_off", c_i64),
        ("rt_off", c_i64),
    ]


class TrsmDesc(ctypes.Structure):
    _fields_ = [
        ("m", c_i64),
        ("l", c_i64),
        ("y_off", c_i64),
        ("y_rs", c_i64),
        ("y_cs", c_i64),
        ("r_off", c_i64),
        ("o_off", c_i64),
    ]


class Gather2dDesc(ctypes.Structure):
    _fields_ = [
        (n, c_i64)
        for n in ("rows", "cols", "src_rows", "src_cols", "s_off", "s_rs", "s_cs", "d_off", "d_rs", "d_cs",
                  "ri_off", "rw_off", "ci_off", "cw_off")
    ] + [("beta", c_dbl)]


class Copy2dDesc(ctypes.Structure):
    _fields_ = [
        ("rows", c_i64),
        ("cols", c_i64),
        ("s_off", c_i64),
        ("s_rs", c_i64),
        ("s_cs", c_i64),
        ("d_off", c_i64),
        ("d_rs", c_i64),
        ("d_cs", c_i64),
        ("rw_off", c_i64),
        ("cw_off", c_i64),
        ("ci_off", c_i64),
        ("alpha", c_dbl),
        ("beta", c_dbl),
    ]


# (name, restype, argtypes) of every exported entry point
_SIGNATURES = {
    "fs_last_error": (ctypes.c_char_p, []),
    "fs_abi_version": (c_int, []),
    "fs_device_count": (c_int, [ctypes.POINTER(c_int)]),
    "fs_u1_fuse": (c_int, [c_int, ctypes.POINTER(c_i8p), c_i64p, c_u8p, c_i8p]),
    "fs_find_block_cuts": (
        c_int,
        [c_int, ctypes.POINTER(c_dblp), c_i64p, c_i64, c_i64p, c_dbl, c_dbl, c_i64p],
    ),
    "fs_abelian_permute_plan_create": (
        c_int,
        [
            c_int,
            c_i64p,
            ctypes.POINTER(c_i8p),
            c_u8p,
            c_int,
            c_int,
            c_i8p,
            c_i64p,
            c_int,
            c_int,
            c_i32p,
            c_int,
            c_int,
            ctypes.POINTER(c_vp),
        ],
    ),
    "fs_permute_plan_info": (c_int, [c_vp, c_i64p, c_i64p, c_i64p]),
    "fs_permute_plan_blocks": (c_int, [c_vp, c_i8p, c_i64p]),
    "fs_permute_plan_tiles": (c_int, [c_vp, c_i64p, c_i64p, c_i64p]),
    "fs_permute_plan_execute_f64": (c_int, [c_vp, c_vp, c_vp, c_vp]),
    "fs_permute_plan_execute_f32": (c_int, [c_vp, c_vp, c_vp, c_vp]),
    "fs_permute_plan_destroy": (c_int, [c_vp]),
    "fs_lie_plan_create": (
        c_int,
        [c_i64, c_i64p, c_i64, c_i64p, c_i64, c_i64p, c_i64, c_dblp, c_i64, c_i64, ctypes.POINTER(c_vp)],
    ),
    "fs_lie_plan_check": (
        c_int, [c_i64, c_i64p, c_i64, c_i64p, c_i64, c_i64p, c_i64, c_i64, c_i64],
    ),
    "fs_lie_plan_tables": (
        c_int,
        [ctypes.POINTER(LiePattern), c_i64, c_i64p, c_i64p, c_i64, c_i64p, c_i64p, c_i64p,
         c_i64p, c_i64, c_i64p, c_i64p, c_i64p],
    ),
    "fs_lie_weights_create": (c_int, [ctypes.POINTER(LiePattern), c_i64, c_dblp, ctypes.POINTER(c_vp)]),
    "fs_lie_weights_destroy": (c_int, [c_vp]),
    "fs_lie_plan_build": (
        c_int,
        [ctypes.POINTER(LiePattern), c_i64, c_i64p, c_i64p, c_i64, c_i64p, c_i64p, c_i64p,
         c_vp, c_i64, c_i64p, ctypes.POINTER(c_vp)],
    ),
    "fs_lie_plan_info": (c_int, [c_vp, c_i64p, c_dblp, c_dblp]),
    "fs_lie_plan_execute_f64": (c_int, [c_vp, c_vp, c_vp, c_vp]),
    "fs_lie_plan_destroy": (c_int, [c_vp]),
    "fs_grouped_gemm_f64": (
        c_int,
        [c_int, ctypes.POINTER(GemmGroup), c_vp, c_vp, c_vp, c_dbl, c_dbl, c_vp],
    ),
    "fs_grouped_gemm_f64_multi": (
        c_int,
        [c_int, ctypes.POINTER(GemmGroup), ctypes.POINTER(c_vp), ctypes.POINTER(c_vp), ctypes.POINTER(c_vp),
         c_dbl, c_dbl, c_vp],
    ),
    "fs_grouped_gemm_f64_noalloc": (
        c_int,
        [c_int, ctypes.POINTER(GemmGroup), c_vp, c_vp, c_vp, c_dbl, c_dbl, c_vp],
    ),
    "fs_gemm_last_dispatches": (c_int, []),
    "fs_grouped_gemm_f64_split": (
        c_int,
        [c_int, ctypes.POINTER(GemmGroup), c_vp, c_vp, c_vp, c_dbl, c_dbl, c_vp],
    ),
    "fs_batched_svd_f64": (
        c_int,
        [c_int, ctypes.POINTER(SvdDesc), c_vp, c_vp, c_vp, c_vp, c_vp, c_int, c_vp],
    ),
    "fs_batched_qr_f64": (
        c_int,
        [c_int, ctypes.POINTER(QrDesc), c_vp, c_vp, c_vp, c_vp, c_vp],
    ),
    "fs_batched_cholesky_f64": (
        c_int,
        [c_int, ctypes.POINTER(CholDesc), c_vp, c_vp, c_vp, c_vp],
    ),
    "fs_batched_triinv_f64": (c_int, [c_int, ctypes.POINTER(TriInvDesc), c_vp, c_vp, c_vp, c_vp]),
    "fs_cholqr_work_size": (c_int, [c_int, ctypes.POINTER(CholQrDesc), c_int, c_int, ctypes.POINTER(c_i64)]),
    "fs_cholqr_f64": (
        c_int,
        [c_int, ctypes.POINTER(CholQrDesc), c_vp, c_vp, c_vp, c_vp, c_i64, c_int, c_int, c_vp, c_vp],
    ),
    "fs_batched_cholinv_f64": (
        c_int,
        [c_int, ctypes.POINTER(CholInvDesc), c_vp, c_vp, c_vp, c_vp, c_vp],
    ),
    "fs_batched_trsm_right_upper_f64": (
        c_int,
        [c_int, ctypes.POINTER(TrsmDesc), c_vp, c_vp, c_vp, c_vp],
    ),
    "fs_weighted_norm2_f64": (c_int, [c_int, c_i64p, c_dblp, c_vp, c_vp, c_vp, c_vp]),
    "fs_scale_f64": (c_int, [c_i64, c_dbl, c_vp, c_vp]),
    "fs_memcpy_async": (c_int, [c_vp, c_vp, c_i64, c_vp]),
    "fs_plan_tables_settle": (c_int, []),
    "fs_scale_rsqrt_dev_f64": (c_int, [c_i64, c_vp, c_vp, c_vp]),
    "fs_axpby_f64": (c_int, [c_i64, c_dbl, c_vp, c_dbl, c_vp, c_vp]),
    "fs_batched_copy2d_f64": (
        c_int,
        [c_int, ctypes.POINTER(Copy2dDesc), c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    ),
    "fs_gather2d_check": (
        c_int, [c_int, ctypes.POINTER(Gather2dDesc), c_i64p, c_i64, c_i64, c_i64, c_i64],
    ),
    "fs_batched_gather2d_f64": (
        c_int, [c_int, ctypes.POINTER(Gather2dDesc), c_vp, c_vp, c_vp, c_vp, c_vp],
    ),
    "fs_weighted_trace_f64": (
        c_int,
        [c_int, c_i64p, c_i64p, c_i64p, c_dblp, c_vp, c_vp, c_vp],
    ),
    "fs_fill_normal_f64": (c_int, [c_i64, ctypes.c_uint64, c_vp, c_vp]),
    "fs_convert_f64_to_f32": (c_int, [c_i64, c_vp, c_vp, c_vp]),
    "fs_convert_f32_to_f64": (c_int, [c_i64, c_vp, c_vp, c_vp]),
}

EXPORTED_SYMBOLS = tuple(_SIGNATURES)


def _load():
    if not _LIB_PATH.exists():
        msg = (
            f"{_LIB_PATH.name} is not built: run `python frostspin_amd/csrc/build.py` "
            "(frostspin_amd has no host fallback for device operations)"
        )
        raise ImportError(msg)
    lib = ctypes.CDLL(os.fspath(_LIB_PATH))
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        # pointer parameters are declared as void pointers: a plain address
        # (int) then passes without building a typed pointer object, while
        # typed pointers, ctypes arrays, byref() and None still pass as before
        fn.argtypes = [c_vp if isinstance(a, type) and issubclass(a, ctypes._Pointer) else a for a in args]
    return lib


lib = _load()


class FrostspinError(RuntimeError):
    """Error reported by the native library."""


def check(status):
    if status != 0:
        msg = lib.fs_last_error().decode(errors="replace")
        raise FrostspinError(msg or f"frostspin_amd error status {status}")


# ----------------------------------------------------------------------------
# small helpers to pass numpy arrays
# ----------------------------------------------------------------------------
# Addresses of contiguous numpy arrays for the void-pointer parameters of the
# entry points. An address does not keep its array alive: the caller must hold
# the array (a name, an attribute, a list element) until the call returns. A
# temporary passed straight in (`i64p(x.copy())`) would be freed before the
# library reads it. Its reference count gives it away: inside the helper only
# the call's own references remain, as many as _TEMP_REFS measures on a
# temporary once at import; such an argument is refused.
def _refs(a):
    return sys.getrefcount(a)


def _refs_via(a):  # the helpers' own call depth: i64p -> _held
    return _refs(a)


_TEMP_REFS = _refs(np.empty(0))
_TEMP_REFS_VIA = _refs_via(np.empty(0))


def _held(a, dtype):
    if sys.getrefcount(a) <= _TEMP_REFS_VIA:
        msg = "array passed to the library by address without a reference held by the caller"
        raise RuntimeError(msg)
    if a.dtype != dtype or not a.flags.c_contiguous:
        msg = f"library argument must be a contiguous {np.dtype(dtype)} array, got {a.dtype}"
        raise TypeError(msg)


def _addr(a):
    if sys.getrefcount(a) <= _TEMP_REFS:
        msg = "array passed to the library by address without a reference held by the caller"
        raise RuntimeError(msg)
    return a.__array_interface__["data"][0]


def i64p(a):
    _held(a, np.int64)
    return a.__array_interface__["data"][0]


def i32p(a):
    _held(a, np.int32)
    return a.__array_interface__["data"][0]


def i8p(a):
    _held(a, np.int8)
    return a.__array_interface__["data"][0]


def u8p(a):
    _held(a, np.uint8)
    return a.__array_interface__["data"][0]


def dblp(a):
    _held(a, np.float64)
    return a.__array_interface__["data"][0]


def device_count():
    n = c_int(0)
    check(lib.fs_device_count(ctypes.byref(n)))
    return n.value


def u1_fuse(reps, flips):
    """Row-major Kronecker sum of int8 charges (int8 wrap-around)."""
    reps = [np.ascontiguousarray(r, dtype=np.int8) for r in reps]
    sizes = np.array([r.size for r in reps], dtype=np.int64)
    flips = np.ascontiguousarray(flips, dtype=np.uint8)
    out = np.empty((int(np.prod(sizes)),), dtype=np.int8)
    ptrs = (c_vp * len(reps))(*[i8p(r) for r in reps])
    check(lib.fs_u1_fuse(len(reps), ptrs, i64p(sizes), u8p(flips), i8p(out)))
    return out


def find_block_cuts_native(block_values, cut_target, dims, cutoff, degen_ratio):
    vals = [np.ascontiguousarray(v, dtype=np.float64) for v in block_values]
    n = len(vals)
    sizes = np.array([v.size for v in vals], dtype=np.int64)
    dims = np.ascontiguousarray(dims, dtype=np.int64)
    cuts = np.zeros((n,), dtype=np.int64)
    if n == 0:
        return cuts
    ptrs = (c_vp * n)(*[dblp(v) for v in vals])
    check(
        lib.fs_find_block_cuts(
            n, ptrs, i64p(sizes), int(cut_target), i64p(dims), float(cutoff),
            float(degen_ratio), i64p(cuts),
        )
    )
    return cuts
