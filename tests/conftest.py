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
