"""Persisted irreps-only structures of non-abelian permutes.

The reference keeps the structural data and the simple-block unitaries of every
permute pattern in process-wide dictionaries
(lie_group_symmetric_tensor.py:432-733, su2_symmetric_tensor.py:175-176), so
every process recomputes them. Here the irreps-only permute structure of a
pattern (task tables and structural matrices, everything that does not depend
on the degeneracies) can be written to a directory shared by several
processes, e.g. the worker slots of a serving engine: a pattern met by one
process is then loaded, not recomputed, by the others and by later runs.

Directory: set_plan_dir(path), or the FROSTSPIN_AMD_PLAN_DIR environment
variable; None (the default) keeps structures in memory only. One .npz per
pattern, named by the SHA-256 of the pattern key and of a format version;
files are written under a temporary name in the same directory and renamed,
so a reader never sees a partial file, and are read with numpy.load without
pickles.
"""

import hashlib
import os
import tempfile
import threading
import time

import numpy as np

FORMAT = "permute-structure-v1"
_STATE = {"dir": os.environ.get("FROSTSPIN_AMD_PLAN_DIR") or None}
_LOCK = threading.Lock()
# structures built / loaded / written by this process and the seconds spent
STATS = {"built": 0, "loaded": 0, "stored": 0, "build_s": 0.0, "load_s": 0.0}


def set_plan_dir(path):
    """Persist structures under path (created if needed); None disables."""
    if path is not None:
        os.makedirs(path, exist_ok=True)
    _STATE["dir"] = None if path is None else os.fspath(path)


def plan_dir():
    return _STATE["dir"]


def pattern_key(symmetry, irreps_per_leg, signature, nrr, axes, nrr_out):
    """Text key of one permute pattern (irreps of every leg, signature, row
    legs before and after, permutation)."""
    legs = ";".join(",".join(str(int(x)) for x in ir) for ir in irreps_per_leg)
    sig = "".join("1" if s else "0" for s in signature)
    ax = ",".join(str(int(a)) for a in axes)
    return f"{FORMAT}|{symmetry}|{legs}|{sig}|{int(nrr)}|{ax}|{int(nrr_out)}"


def _path(key):
    return os.path.join(_STATE["dir"], hashlib.sha256(key.encode()).hexdigest()[:40] + ".npz")


def load(key):
    """Arrays stored for key, or None (disabled, absent, or another key hashed
    to the same name)."""
    if _STATE["dir"] is None:
        return None
    path = _path(key)
    if not os.path.exists(path):
        return None
    t0 = time.perf_counter()
    try:
        with np.load(path, allow_pickle=False) as fin:
            if str(fin["_key"]) != key:
                return None
            arrays = {k: fin[k] for k in fin.files if k != "_key"}
    except (OSError, ValueError, KeyError):
        return None  # unreadable file: rebuilt and rewritten by the caller
    with _LOCK:
        STATS["loaded"] += 1
        STATS["load_s"] += time.perf_counter() - t0
    return arrays


def store(key, arrays):
    """Write arrays for key (atomically); no-op when disabled."""
    if _STATE["dir"] is None:
        return
    fd, tmp = tempfile.mkstemp(dir=_STATE["dir"], suffix=".tmp")
    try:
        with os.fdopen(fd, "wb") as fout:
            np.savez(fout, _key=np.array(key), **arrays)
        os.replace(tmp, _path(key))
    except OSError:
        if os.path.exists(tmp):
            os.remove(tmp)
        return
    with _LOCK:
        STATS["stored"] += 1


def note_built(seconds):
    with _LOCK:
        STATS["built"] += 1
        STATS["build_s"] += seconds


def stats():
    with _LOCK:
        return dict(STATS)
