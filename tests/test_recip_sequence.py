"""The batched 2D force pass computes 1/r without IEEE division
(recip_rn, csrc/cwa_batch2d.hip): a hardware reciprocal, two Newton-Raphson
steps and a carry fix-up.  These host checks run the same steps from every
start within 3 ulp of the answer and demand IEEE 1.0f / b bit for bit, over
every mantissa of several binades (the steps scale exactly by powers of two
in [2^-40, 2^40]).  They also show the fix-up is needed: without it the
all-ones mantissas fail."""
import shutil
import subprocess

import pytest

from pathlib import Path

SRC = Path(__file__).parent / "cpp" / "recip_rn_check.c"


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not found")
    exe = tmp_path_factory.mktemp("recip") / "recip_rn_check"
    # no contraction: the fmaf calls are the only fused operations
    subprocess.run([gcc, "-O2", "-ffp-contract=off", "-o", str(exe), str(SRC), "-lm"], check=True)
    return exe


@pytest.mark.parametrize("binades", [(-1, 1), (-40, -40), (40, 40)])
def test_reciprocal_steps_are_correctly_rounded(checker, binades):
    out = subprocess.run([str(checker), str(binades[0]), str(binades[1]), "3"],
                         capture_output=True, text=True, timeout=300)
    checked, bad, bad_unfixed = map(int, out.stdout.split())
    nb = binades[1] - binades[0] + 1
    assert checked == nb * (1 << 23) * 7
    assert bad == 0 and out.returncode == 0
    # the all-ones mantissa of each binade fails without the fix-up
    assert bad_unfixed > 0
