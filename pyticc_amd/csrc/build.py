"""Build libticc_hip.so for gfx950 with hipcc (in-tree, no JIT cache).

Usage: python -m pyticc_amd.csrc.build   (or pyticc_amd.csrc.build.build())
The library lands in pyticc_amd/_lib/libticc_hip.so.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
PKG = HERE.parent
REPO = PKG.parent
OUT_DIR = PKG / "_lib"
LIB = OUT_DIR / "libticc_hip.so"
SOURCES = ("util.hip", "gj.hip", "gemm.hip", "ldmd.hip", "contract.hip", "match.hip", "mlp.hip", "noalloc.hip", "capi.hip")
ARCH = os.environ.get("TICC_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    # packed FP32 instructions stay disabled for every kernel of this library
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    # the Gauss-Jordan panel keeps its rows in registers only if its column
    # loop is fully unrolled (otherwise the row arrays move to scratch)
    "-mllvm",
    "-pragma-unroll-threshold=200000",
    "-Wno-unused-command-line-argument",
    f"-I{REPO / 'include'}",
    f"-I{HERE}",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the HIP backend cannot be built")


def build(verbose: bool = False) -> Path:
    """Compile every source to an object, then link the shared library."""
    hipcc = _hipcc()
    obj_dir = OUT_DIR / "obj"
    obj_dir.mkdir(parents=True, exist_ok=True)

    def compile_one(name: str) -> Path:
        src = HERE / name
        obj = obj_dir / (name.replace(".hip", ".o"))
        cmd = [hipcc, *FLAGS, "-c", "-x", "hip", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"hipcc failed for {name}:\n{res.stdout}\n{res.stderr}")
        return obj

    workers = min(len(SOURCES), max(1, int(os.environ.get("MAX_JOBS", "4"))))
    with ThreadPoolExecutor(max_workers=workers) as pool:
        objs = list(pool.map(compile_one, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv)
    print(path)
