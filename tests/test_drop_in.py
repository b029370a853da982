"""Drop-in surface that reference scripts touch directly.

* `frostspin_amd.config` (frostspin/config.py:15-28): the SU(2)
  Clebsch-Gordan table option `--frostspin-SU2-CG-file` is read from the
  command line at import; a table named there replaces the computed
  coefficients, a missing file raises FileNotFoundError like the reference.
* `blocks` returns host numpy arrays (symmetric_tensor.py:320); writing into
  them leaves the tensor unchanged, `device_blocks` are the device views.
* `expm` of general (non-symmetric) blocks against scipy.linalg.expm per
  block, the reference's own implementation (symmetric_tensor.py:818-824):
  missing blocks stay missing, a zero block gives the identity.
"""

import os
import subprocess
import sys

import numpy as np
import pytest
import scipy.linalg as lg

from frostspin_amd import SU2SymmetricTensor, U1SymmetricTensor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

BACKENDS = [
    pytest.param("emulated", id="emulated"),
    pytest.param("gpu", id="gpu", marks=pytest.mark.gpu),
]


@pytest.fixture(params=BACKENDS)
def backend(request):
    if request.param == "emulated":
        request.getfixturevalue("emulated_device")
    else:
        request.getfixturevalue("gpu")
    return request.param


def _run(code, *args):
    return subprocess.run(
        [sys.executable, "-c", code, *args], cwd=ROOT, capture_output=True, text=True, timeout=120
    )


def test_config_exports_and_cg_file_option(tmp_path):
    import frostspin_amd

    assert set(frostspin_amd.config) >= {"SU2_CG_file", "debug_checks"}
    assert frostspin_amd.ASSERT_TOL == 4e-13
    assert "config" in frostspin_amd.__all__
    # a table written in the reference's JSON format, with one tensor altered
    # so that its use is visible
    table = tmp_path / "cg.json"
    code_write = (
        "import json, sys; from frostspin_amd.groups import su2_clebsch_gordan as cg;"
        "cg.save_su2_cg(sys.argv[1], 3);"
        "d = json.load(open(sys.argv[1])); d['2,2,1'][1] = [2 * x for x in d['2,2,1'][1]];"
        "json.dump(d, open(sys.argv[1], 'w'))"
    )
    r = _run(code_write, str(table))
    assert r.returncode == 0, r.stderr
    code_read = (
        "import sys, numpy as np; import frostspin_amd;"
        "from frostspin_amd.groups import su2_clebsch_gordan as cg;"
        "print(frostspin_amd.config['SU2_CG_file']);"
        "print(float(np.abs(cg.clebsch_gordan(2, 2, 1)).max()), float(np.abs(cg.clebsch_gordan(3, 3, 1)).max()))"
    )
    r = _run(code_read, "--frostspin-SU2-CG-file", str(table))
    assert r.returncode == 0, r.stderr
    lines = r.stdout.split("\n")
    assert lines[0] == str(table)
    big, other = (float(x) for x in lines[1].split())
    assert abs(big - 2 * np.sqrt(0.5)) < 1e-15  # from the altered table
    assert abs(other - np.sqrt(1 / 3)) < 1e-15  # d = 3 computed (outside the table)
    r = _run(code_read)  # no option: computed coefficients
    assert r.returncode == 0, r.stderr
    assert r.stdout.split("\n")[0] == "None"
    assert abs(float(r.stdout.split("\n")[1].split()[0]) - np.sqrt(0.5)) < 1e-15
    r = _run("import frostspin_amd", "--frostspin-SU2-CG-file", str(tmp_path / "missing.json"))
    assert r.returncode != 0 and "FileNotFoundError" in r.stderr


def test_blocks_are_host_arrays(backend):
    rng = np.random.default_rng(3)
    leg = np.array([-1, 0, 0, 1], dtype=np.int8)
    t = U1SymmetricTensor.random((leg, leg), (leg, leg), rng=rng)
    blocks = t.blocks
    assert all(isinstance(b, np.ndarray) and b.dtype == np.float64 for b in blocks)
    dense = t.toarray()
    blocks[0][:] = 7.0  # a copy: the tensor does not change
    assert (t.toarray() == dense).all()
    tt = t.transpose()  # column-major blocks come back with their logical shape
    by_irrep = dict(zip((-tt.block_irreps).tolist(), tt.blocks, strict=True))  # U(1): conjugate irrep
    for irr, b in zip(t.block_irreps.tolist(), t.blocks, strict=True):
        assert (b.T == by_irrep[irr]).all()
    dev = t.device_blocks
    assert all(tuple(d.shape) == b.shape for d, b in zip(dev, t.blocks, strict=True))


def _check_expm(t, tol):
    e = t.expm()
    assert (e.block_irreps == t.block_irreps).all()
    for b, eb in zip(t.blocks, e.blocks, strict=True):
        ref = lg.expm(b)
        assert np.abs(eb - ref).max() <= tol * max(1.0, np.abs(ref).max())
    return e


def test_expm_general_blocks(backend):
    rng = np.random.default_rng(11)
    leg = np.array([-2, 0, 0, 2, 2], dtype=np.int8)
    t = U1SymmetricTensor.random((leg, leg), (leg, leg), rng=rng)
    _check_expm(t, 1e-13)  # small norm, no squaring
    _check_expm(t * 9.0, 1e-12)  # several squarings
    # symmetric blocks, like the simple-update gates
    h = U1SymmetricTensor(t.row_reps, t.col_reps, [b + b.T for b in t.blocks], t.block_irreps, t.signature)
    _check_expm(-0.01 * h, 1e-14)
    # a zero block exponentiates to the identity, a missing block stays missing
    blocks = list(t.blocks)
    blocks[0] = np.zeros_like(blocks[0])
    tz = U1SymmetricTensor(t.row_reps, t.col_reps, blocks[:-1], t.block_irreps[:-1], t.signature)
    ez = _check_expm(tz, 1e-13)
    assert (ez.blocks[0] == np.eye(blocks[0].shape[0])).all()
    assert ez.nblocks == t.nblocks - 1
    # SU(2): blocks of a random operator on spin-1/2 x spin-1 legs
    r = np.array([[1, 2], [2, 3]])
    s = SU2SymmetricTensor.random((r, r), (r, r), rng=rng)
    _check_expm(s, 1e-13)
    # the reference's tensor-level identity: exp of a block-diagonal operator
    # equals the matrix exponential of its dense form
    dense = s.toarray(as_matrix=True)
    assert np.abs(s.expm().toarray(as_matrix=True) - lg.expm(dense)).max() < 1e-12 * np.abs(lg.expm(dense)).max()
