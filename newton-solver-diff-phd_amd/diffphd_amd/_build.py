"""Build recipe of the native library ``libdiffphd.so`` (gfx950).

Host setup sources (``factor.cpp``, ``layout.cpp``, ``capi_setup.cpp``) are compiled by g++
with ``-ffp-contract=off`` so their arithmetic rounds where the reference's
NumPy rounds; HIP sources are compiled by hipcc for gfx950 only (the contact
kernels also without contraction). The library
is written in-tree next to this file.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
ROOT = PACKAGE.parent
CSRC = ROOT / "csrc"
INCLUDE = ROOT.parent / "include"
BUILD = ROOT / "_build"
LIBRARY = PACKAGE / "libdiffphd.so"
ARCH = os.environ.get("DPHD_OFFLOAD_ARCH", "gfx950")

HOST_SOURCES = ["factor.cpp", "layout.cpp", "capi_setup.cpp"]
HIP_SOURCES = ["kernels.hip", "contact.hip", "solver.hip"]
# Contact arithmetic (closest points, pair forces, the frictional Newton step) rounds as the
# reference's does: no fused multiply-adds anywhere in that file, the shared vector helpers included.
HIP_FLAGS = {"contact.hip": ["-ffp-contract=off"]}


def _hipcc():
    for candidate in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if candidate and Path(candidate).exists():
            return candidate
    raise RuntimeError("hipcc not found: the native library cannot be built")


def _compile(cmd):
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{proc.stdout}\n{proc.stderr}")
    return proc


def build(verbose: bool = False) -> Path:
    """Compile every source and link ``libdiffphd.so``; always recompiles."""
    BUILD.mkdir(exist_ok=True)
    hipcc = _hipcc()
    jobs = []
    objects = []
    for src in HOST_SOURCES:
        obj = BUILD / (src + ".o")
        objects.append(obj)
        jobs.append(
            [
                "g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra",
                f"-I{INCLUDE}", "-c", str(CSRC / src), "-o", str(obj),
            ]
        )
    for src in HIP_SOURCES:
        obj = BUILD / (src + ".o")
        objects.append(obj)
        jobs.append(
            [
                hipcc, "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
                "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", *HIP_FLAGS.get(src, []),
                "-Wall", f"-I{INCLUDE}", "-c", str(CSRC / src), "-o", str(obj),
            ]
        )
    with ThreadPoolExecutor(max_workers=min(4, os.cpu_count() or 1)) as pool:
        for proc in pool.map(_compile, jobs):
            if verbose and proc.stderr.strip():
                print(proc.stderr, file=sys.stderr)
    tmp = LIBRARY.with_suffix(f".so.tmp{os.getpid()}")
    _compile([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objects), "-o", str(tmp)])
    os.replace(tmp, LIBRARY)
    return LIBRARY


if __name__ == "__main__":
    print(build(verbose=True))
