"""The batched 2D pair loops compute 1/r and sqrt(r^2) without the library
sequences (recip_rn, sqrt_rn_mid, sqrt_rn_clamped in csrc/cwa_batch2d.hip): a
hardware estimate refined by Newton-Raphson steps in fused multiply-adds.
These host checks run the same steps from every start within 3 ulp of the
answer and demand IEEE 1.0f / b and sqrtf(x) bit for bit, over every float
of several binades (the steps scale exactly by powers of two in the ranges
the kernels use them in).  The reciprocal check also shows its fix-up is
needed: without it the all-ones mantissas fail.  The last checks pin the
float facts that let the pair loops drop branches of the shader's kernel
functions without changing a bit: W is flat below the square root's clamp,
and q = r (1/H) stays at most 1 where the q >= 1 early returns were."""
import shutil
import subprocess

import pytest

from pathlib import Path

CPP = Path(__file__).parent / "cpp"


def _build(tmp_path_factory, name):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not found")
    exe = tmp_path_factory.mktemp(name) / name
    # no contraction: the fmaf calls are the only fused operations
    subprocess.run([gcc, "-O2", "-ffp-contract=off", "-o", str(exe), str(CPP / f"{name}.c"), "-lm"],
                   check=True)
    return exe


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    return _build(tmp_path_factory, "recip_rn_check")


@pytest.fixture(scope="module")
def sqrt_checker(tmp_path_factory):
    return _build(tmp_path_factory, "sqrt_rn_check")


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


@pytest.mark.parametrize("binades", [(0, 2), (-80, -78), (78, 80)])
def test_square_root_steps_are_correctly_rounded(sqrt_checker, binades):
    out = subprocess.run([str(sqrt_checker), str(binades[0]), str(binades[1]), "3"],
                         capture_output=True, text=True, timeout=300)
    checked, bad, zero_ok, max_seed = map(int, out.stdout.split())
    assert checked == (binades[1] - binades[0]) * (1 << 23) * 7
    assert bad == 0 and zero_ok == 1 and out.returncode == 0
    # 2h seeds the reciprocal steps, which are checked from starts within 3 ulp
    assert max_seed <= 3


@pytest.mark.parametrize("h", [2.0 ** -16, 2.0 ** -16 * 1.7, 0.02, 0.3, 1.0, 2.0 ** 19])
def test_density_kernel_flat_below_the_square_root_clamp(h):
    """The batched density pass clamps r^2 at 2^-80 (r = 2^-40) before its
    square root.  That is exact because W(r) rounds to k2 for every
    r <= 2^-40 when 2^-16 <= H <= 2^19: the same float operations as the
    kernel's first piece, k2 (6 q q (q - 1) + 1) with q = r (1/H)."""
    import numpy as np

    f = np.float32
    H = f(h)
    HSQ = H * H
    invH = f(1) / H
    k2 = f(40) / (f(7) * f(3.14159265) * HSQ)
    tiny = np.array([0.0, 1e-45, 1e-40, 2.0 ** -100, 2.0 ** -60, 2.0 ** -40.5,
                     np.nextafter(f(2.0 ** -40), f(0)), 2.0 ** -40], np.float32)
    for r in tiny:
        q = r * invH
        assert q <= f(0.5)
        w = k2 * (f(6) * q * q * (q - f(1)) + f(1))
        assert w == k2, (h, r, w, k2)


def test_kernel_gradient_argument_stays_at_most_one():
    """The batched force pass evaluates W'(r) only for r < H and drops the
    q >= 1 early return: q = r * (1/H) with both rounded to float is then at
    most 1.  Checked for the largest float below H (q grows with r) over
    every mantissa of one binade and random H across [2^-16, 2^19]."""
    import numpy as np

    m = np.arange(1 << 23, dtype=np.uint32)
    H = (np.uint32(127 << 23) | m).view(np.float32)           # [1, 2)
    rng = np.random.default_rng(0)
    H = np.concatenate([H, rng.uniform(2.0 ** -16, 2.0 ** 19, 1 << 20).astype(np.float32),
                        np.float32([2.0 ** -16, 2.0 ** 19])])
    r = np.nextafter(H, np.float32(0))
    q = r * (np.float32(1) / H)
    assert q.dtype == np.float32 and float(q.max()) <= 1.0


def test_kernel_argument_of_the_density_test_stays_at_most_one():
    """The batched density pass evaluates W(sqrt(r2)) only for r2 < HSQ (the
    shader's test) and drops W's q >= 1 early return: the largest float
    below HSQ = H*H, its correctly rounded square root times the rounded
    1/H, is at most 1 (q grows with r2), and at q == 1 the second piece
    k2x2 (1-q)^3 is +0 like the early return.  Every mantissa of one
    binade of H (q does not change when H scales by a power of two)."""
    import numpy as np

    m = np.arange(1 << 23, dtype=np.uint32)
    H = (np.uint32(127 << 23) | m).view(np.float32)
    r2 = np.nextafter(H * H, np.float32(0))
    q = np.sqrt(r2) * (np.float32(1) / H)
    assert q.dtype == np.float32 and float(q.max()) <= 1.0
    q1 = np.float32(1) - np.float32(1)
    assert (np.float32(3.0) * q1 * q1 * q1) == 0 and not np.signbit(np.float32(3.0) * q1 * q1 * q1)
