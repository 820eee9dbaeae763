"""The C ABI from a C++ host (tests/cpp/cwa_consumer.cpp): the reference
application's own scene built and stepped through include/cwa.h alone.

CPU: the program builds and links against libcwa.so, and its device-free
checks pass (ABI version and struct sizes as C++ sees them, reference
parameters, an invalid scene rejected before any device work).
GPU: 20 steps of the reference scene run by the C++ host equal, bit for
bit, the same calls made through the Python binding in this process (the
step is deterministic: no float atomics).
"""
import runpy
import subprocess
from pathlib import Path

import numpy as np
import pytest

HERE = Path(__file__).resolve().parent
BIN = HERE / "cpp" / "_build" / "cwa_consumer"


def _binary():
    if not BIN.exists():
        runpy.run_path(str(HERE / "cpp" / "build.py"))["build"](verbose=False)
    return BIN


def test_consumer_builds_and_checks(cwa):
    r = subprocess.run([str(_binary()), "check"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert "invalid scene -> -1" in r.stdout


@pytest.mark.gpu
def test_consumer_run_matches_python_binding(gpu, cwa, tmp_path):
    from coupledwateranimation_amd import configs

    steps = 20
    out = tmp_path / "state.bin"
    print("cpp consumer: reference scene,", steps, "steps", flush=True)
    r = subprocess.run([str(_binary()), "run", str(steps), str(out)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    n, W, H = 20480, 64, 64
    raw = np.fromfile(out, np.float32)
    assert raw.size == 4 * n + 4 * n + n + n + W * H
    pos, vel = raw[:4 * n].reshape(n, 4), raw[4 * n:8 * n].reshape(n, 4)
    rho, prs = raw[8 * n:9 * n], raw[9 * n:10 * n]
    h0 = raw[10 * n:].reshape(H, W)
    spec = configs.reference_scene()
    s = cwa.Scene(spec.params, spec.pos, None, None, (W, H))
    s.wave_reinit()
    s.step(5e-5, steps)
    st = s.readback()
    s.close()
    for name, a, b in (("pos", pos, st["pos"]), ("vel", vel, st["vel"]), ("rho", rho, st["rho"]),
                       ("prs", prs, st["prs"]), ("height0", h0, st["height0"])):
        assert np.array_equal(np.asarray(a, np.float32).view(np.uint32),
                              np.asarray(b, np.float32).reshape(a.shape).view(np.uint32)), name
