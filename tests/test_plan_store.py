"""Persisted permute structures (plan_store): a structure written by one
process is loaded by another and gives the same permute; the file holds only
arrays (no pickles) under a name derived from the pattern key; another key
never loads it."""

import os
import subprocess
import sys

import numpy as np
import pytest

from frostspin_amd import SU2SymmetricTensor
from frostspin_amd.symmetric_tensor import plan_store
from frostspin_amd.symmetric_tensor.lie_group_symmetric_tensor import LieGroupSymmetricTensor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import sys
sys.path.insert(0, {root!r})
import numpy as np
from oracle import device_emulator
device_emulator.install_global()
from frostspin_amd import SU2SymmetricTensor
from frostspin_amd.symmetric_tensor import plan_store
plan_store.set_plan_dir({d!r})
r = np.array([[1, 2, 1], [1, 2, 3]])
t = SU2SymmetricTensor.random((r, r), (r, r), rng=np.random.default_rng(0))
np.save({out!r}, t.permute((1, 3), (0, 2)).toarray())
s = plan_store.stats()
print(s["built"], s["loaded"])
"""


def _forget():
    for cache in (LieGroupSymmetricTensor._permute_structures, LieGroupSymmetricTensor._structures_by_bytes,
                  LieGroupSymmetricTensor._plans):
        cache.clear()


@pytest.fixture
def store_dir(tmp_path):
    old = plan_store.plan_dir()
    plan_store.set_plan_dir(tmp_path / "plans")
    _forget()
    yield tmp_path / "plans"
    plan_store.set_plan_dir(old)
    _forget()


def test_structure_shared_between_processes(emulated_device, store_dir, tmp_path):
    r = np.array([[1, 2, 1], [1, 2, 3]])
    t = SU2SymmetricTensor.random((r, r), (r, r), rng=np.random.default_rng(0))
    before = plan_store.stats()
    ref = t.permute((1, 3), (0, 2)).toarray()
    after = plan_store.stats()
    assert after["built"] > before["built"] and after["stored"] > before["stored"]
    files = os.listdir(store_dir)
    assert files and all(f.endswith(".npz") for f in files)
    for f in files:  # arrays only: loads without pickles
        with np.load(store_dir / f, allow_pickle=False) as fin:
            assert str(fin["_key"]).startswith(plan_store.FORMAT)
    out = tmp_path / "child.npy"
    code = CHILD.format(root=ROOT, d=os.fspath(store_dir), out=os.fspath(out))
    res = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, check=True,
                         env={**os.environ, "HIP_VISIBLE_DEVICES": ""})
    built, loaded = (int(x) for x in res.stdout.split()[-2:])
    assert built == 0 and loaded > 0  # every structure came from the directory
    assert np.array_equal(np.load(out), ref)


def test_other_keys_do_not_load(store_dir):
    key = plan_store.pattern_key("SU2", ((1, 3), (2,)), (False, True), 1, (1, 0), 1)
    plan_store.store(key, {"x": np.arange(3)})
    assert np.array_equal(plan_store.load(key)["x"], np.arange(3))
    other = plan_store.pattern_key("SU2", ((1, 3), (2,)), (False, True), 1, (1, 0), 0)
    assert plan_store.load(other) is None
    # a file whose stored key differs (hash clash) is not used
    os.replace(plan_store._path(key), plan_store._path(other))
    assert plan_store.load(other) is None
