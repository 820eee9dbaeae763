"""Shared test setup: import paths, the `gpu` marker, library fixtures."""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); calls the C ABI")


@pytest.fixture(scope="session")
def oracle():
    """The CPU oracle (built on demand)."""
    from oracle import oracle as O

    if not O.LIB_PATH.exists():
        subprocess.run(["make", "-C", str(ROOT / "oracle")], check=True)
    O.lib()
    return O


@pytest.fixture(scope="session")
def cwa():
    """The HIP library binding; fails (never skips) when it is missing."""
    import coupledwateranimation_amd as pkg

    pkg.lib()
    return pkg


@pytest.fixture(scope="session")
def gpu(cwa):
    """A visible HIP device; GPU tests fail loudly without one."""
    n = cwa.lib().cwa_device_count()
    if n < 1:
        pytest.fail("no HIP device visible (GPU tests must run on the MI355X box)")
    import torch

    assert torch.cuda.is_available()
    return torch.device("cuda:0")


@pytest.fixture(params=[4, 1], ids=["four_lanes", "one_lane"])
def lanes(request, cwa):
    """Lanes per particle of the density pass (and of the pair adjoint of
    tapes created in the test): 4 is the form of scenes up to 2^18
    particles, 1 the form of larger scenes, forced here on small ones."""
    from coupledwateranimation_amd.scene import set_lane_split_limit

    prev = set_lane_split_limit(-1 if request.param == 4 else 0)
    yield request.param
    set_lane_split_limit(prev)
