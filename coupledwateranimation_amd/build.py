"""Build recipe of the HIP library (libcwa.so) for gfx950.

The library is built in-tree (coupledwateranimation_amd/_build/libcwa.so)
with hipcc directly; no torch extension machinery is involved, the library
is loaded through ctypes.  ``python -m coupledwateranimation_amd.build``
runs it.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "_build"
LIB = OUT / "libcwa.so"
ROOT = PKG.parent

SOURCES = ["cwa_kernels.hip", "cwa_adjoint.hip", "cwa_scene.hip", "cwa_tape.hip", "cwa_batch2d.hip",
           "cwa_gen.cpp"]
# The batched 2D step reproduces the oracle's float expressions bit for bit:
# no contraction into fused multiply-adds beyond the explicit fmaf calls.
EXTRA = {"cwa_batch2d.hip": ["-ffp-contract=off"]}
HEADERS = ["cwa_device.hpp", "cwa_kernels.hpp", "cwa_adjoint.hpp", "cwa_modules.hpp"]

ARCH = os.environ.get("CWA_OFFLOAD_ARCH", "gfx950")

# Device code: no packed fp32 (v_pk_*_f32) instructions.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
          "-I", str(ROOT / "include")]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = True) -> Path:
    """Compile every source for gfx950 and link libcwa.so (always rebuilds)."""
    OUT.mkdir(parents=True, exist_ok=True)
    cc = hipcc()
    objs = []
    for src in SOURCES:
        obj = OUT / (Path(src).stem + ".o")
        cmd = [cc, *COMMON, "-c", str(CSRC / src), "-o", str(obj)]
        if src.endswith(".hip"):
            cmd[1:1] = ["-x", "hip", f"--offload-arch={ARCH}", *DEVICE_FLAGS]
        cmd[1:1] = EXTRA.get(src, [])
        _run(cmd)
        objs.append(str(obj))
    tmp = LIB.with_suffix(".so.tmp")
    _run([cc, "-shared", f"--offload-arch={ARCH}", "-o", str(tmp), *objs])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
