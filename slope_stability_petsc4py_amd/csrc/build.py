"""Compile the ssx HIP library for gfx950 into the package directory (in-tree, so it travels with
the source tree).  Usage: ``python -m slope_stability_petsc4py_amd.csrc.build``.

Every object is compiled with hipcc --offload-arch=gfx950; packed-FP32 instruction selection is
disabled for all device code.  Objects are rebuilt every call (the build is the build check).
"""

from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
PKG = HERE.parent
OUT = PKG / "libssx.so"
OBJDIR = HERE / "_objs"

SOURCES = [
    "runtime.cpp",
    "constitutive.hip",
    "assembly.hip",
    "sparse.hip",
    "coarse.hip",
    "p4_mfma.hip",
    "mlp.hip",
    "node_assembly.hip",
    "block_assembly.hip",
    "affine_assembly.hip",
    "probe.hip",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = [
    "-O3",
    "-fPIC",
    "-std=c++17",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _compile(src: str) -> Path:
    obj = OBJDIR / (Path(src).stem + ".o")
    cmd = [HIPCC, *COMMON, "-c", str(HERE / src), "-o", str(obj)]
    if src.endswith(".cpp"):
        cmd = [HIPCC, *COMMON, "-x", "hip", "-c", str(HERE / src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = True) -> Path:
    OBJDIR.mkdir(exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
