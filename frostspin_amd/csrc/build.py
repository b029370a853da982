"""Build libfrostspin_amd.so for gfx950 with hipcc (in-tree, no JIT cache).

Usage: python frostspin_amd/csrc/build.py   (or __graft_entry__.build())
The shared library lands next to the package (frostspin_amd/libfrostspin_amd.so)
so it travels with the source tree.
"""

import concurrent.futures
import os
import shutil
import subprocess
import sys
import tempfile
from pathlib import Path

CSRC = Path(__file__).resolve().parent
PKG = CSRC.parent
ROOT = PKG.parent
INCLUDE = ROOT / "include"
LIB_NAME = "libfrostspin_amd.so"
SOURCES = [
    "fs_planner.cpp",
    "fs_cholqr.cpp",
    "fs_permute.hip",
    "fs_gemm.hip",
    "fs_svd.hip",
    "fs_qr.hip",
    "fs_chol.hip",
    "fs_lie.hip",
    "fs_elementwise.hip",
    "fs_runtime.hip",
]
ARCH = os.environ.get("FS_OFFLOAD_ARCH", "gfx950")

# packed-fp32 instructions are disabled for all device code (they have been
# observed to give silently wrong results on this hardware generation).
DEVICE_FLAGS = [
    f"--offload-arch={ARCH}",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]
COMMON_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", f"-I{INCLUDE}"]


def _hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(exe).exists():
        raise RuntimeError("hipcc not found: the ROCm toolchain is required to build")
    return exe


def _compile(src, obj):
    cmd = [_hipcc(), *COMMON_FLAGS]
    if src.suffix == ".hip":
        cmd += ["-x", "hip", *DEVICE_FLAGS]
    else:
        cmd += ["-x", "c++"]
    cmd += ["-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True, check=False)
    if res.returncode != 0:
        msg = f"compilation failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}"
        raise RuntimeError(msg)
    return res.stderr


def build(verbose=False):
    out = PKG / LIB_NAME
    with tempfile.TemporaryDirectory(prefix="fs_build_") as tmp:
        objs = []
        jobs = []
        with concurrent.futures.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
            for s in SOURCES:
                src = CSRC / s
                obj = Path(tmp) / (src.stem + ".o")
                objs.append(obj)
                jobs.append(ex.submit(_compile, src, obj))
            for j in jobs:
                log = j.result()
                if verbose and log.strip():
                    print(log, file=sys.stderr)
        tmp_out = Path(tmp) / LIB_NAME
        cmd = [_hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs)]
        cmd += ["-o", str(tmp_out)]
        res = subprocess.run(cmd, capture_output=True, text=True, check=False)
        if res.returncode != 0:
            raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
        # replace the library by a rename: a process that has the old file
        # mapped keeps its copy (overwriting it in place would fault that process)
        staged = out.with_name(out.name + ".tmp")
        shutil.copy2(tmp_out, staged)
        os.replace(staged, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
