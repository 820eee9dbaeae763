"""Build recipe of the C++ ABI consumer (tests/cpp/_build/cwa_consumer).

It links the in-tree libcwa.so through include/cwa.h alone, as a C++ host
application would; the rpath finds the library relative to the binary, so
the built tree runs wherever it is copied.
"""
from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
SRC = HERE / "cwa_consumer.cpp"
OUT = HERE / "_build" / "cwa_consumer"
LIBDIR = ROOT / "coupledwateranimation_amd" / "_build"
ROCM_LIB = Path(os.environ.get("ROCM_PATH", "/opt/rocm")) / "lib"


def build(verbose: bool = True) -> Path:
    OUT.parent.mkdir(parents=True, exist_ok=True)
    cxx = shutil.which("g++") or "g++"
    rel = os.path.relpath(LIBDIR, OUT.parent)
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-I", str(ROOT / "include"), str(SRC),
           "-L", str(LIBDIR), "-lcwa", "-L", str(ROCM_LIB),
           f"-Wl,-rpath,$ORIGIN/{rel}", f"-Wl,-rpath,{ROCM_LIB}", "-o", str(OUT)]
    if verbose:
        print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return OUT


if __name__ == "__main__":
    build()
