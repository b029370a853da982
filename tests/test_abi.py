"""The C-ABI library loads on a CPU host, exports every entry point declared in
include/frostspin_amd.h, reports argument errors through fs_last_error, and
device entry points refuse to run without device tensors."""

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from frostspin_amd import _lib, _ops

HEADER = Path(__file__).resolve().parents[1] / "include" / "frostspin_amd.h"


def declared_symbols():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fs_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    decl = declared_symbols()
    assert decl, "no declarations parsed"
    assert sorted(_lib.EXPORTED_SYMBOLS) == decl


def test_library_exports_every_symbol():
    lib = ctypes.CDLL(str(_lib._LIB_PATH))
    for name in declared_symbols():
        assert hasattr(lib, name), name
    assert _lib.lib.fs_abi_version() == 1
    assert _lib.device_count() >= 0


def test_planner_errors_are_reported():
    shape = np.array([2, 2], dtype=np.int64)
    legs = [np.zeros(2, np.int8), np.zeros(2, np.int8)]
    with pytest.raises(_lib.FrostspinError, match="axes"):
        _ops.abelian_permute_plan(shape, legs, [0, 1], 1, np.zeros((1,), np.int8), None, False,
                                  [0, 0], 1)
    with pytest.raises(_lib.FrostspinError, match="sorted"):
        _ops.abelian_permute_plan(shape, legs, [0, 1], 1, np.array([0, 0], np.int8), None, False,
                                  [1, 0], 1)
    with pytest.raises(_lib.FrostspinError, match="degen_ratio"):
        _lib.find_block_cuts_native([np.ones(2)], 1, [1], 0.0, 1.5)


def test_no_host_fallback_for_device_ops():
    import torch

    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(RuntimeError, match="HIP GPU"):
        _ops.empty(4)
    with pytest.raises(RuntimeError, match="device tensors"):
        _ops.scale_(torch.ones(3, dtype=torch.float64), 2.0)


def test_address_helpers_refuse_temporaries():
    """Pointer parameters take bare addresses: an array the caller does not
    hold would be freed before the library reads it, so the helpers refuse it;
    held arrays (names, attributes, list elements) pass, wrong dtypes raise."""
    import numpy as np
    import pytest as _pytest

    from frostspin_amd import _lib

    held = np.arange(4, dtype=np.int64)
    assert _lib.i64p(held) == held.__array_interface__["data"][0]
    assert [_lib.dblp(v) for v in [np.zeros(2)]][0] != 0
    with _pytest.raises(RuntimeError):
        _lib.i64p(held.copy())
    with _pytest.raises(RuntimeError):
        _lib._addr(np.zeros(3))
    with _pytest.raises(TypeError):
        _lib.i32p(held)
