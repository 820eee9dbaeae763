"""Host-side checks of the C ABI that need no GPU: the library loads,
exports every function include/cwa.h declares, struct layouts match the
binding, generators are deterministic, invalid parameters are rejected
before any device work."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared_functions(header="cwa.h", prefix="cwa_"):
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(" + prefix + r"[a-z0-9_]+)\s*\(", text)))


def test_exports_every_declared_symbol(cwa):
    L = cwa.lib()
    names = _declared_functions()
    assert len(names) >= 25
    names2d = _declared_functions("cwa2d.h", "cwa2d_")
    assert len(names2d) >= 20
    missing = [n for n in names + names2d if not hasattr(L, n)]
    assert not missing, missing
    from coupledwateranimation_amd import _lib
    assert not [n for n in names + names2d if n not in _lib.SIGNATURES]


def test_2d_struct_sizes_and_rejections(cwa):
    """cwa2d structs match the binding; invalid batches are rejected before
    any device work (no GPU needed)."""
    from coupledwateranimation_amd import _lib, batch2d

    L = cwa.lib()
    assert L.cwa2d_abi_version() == 2
    for k, st in enumerate(_lib.STRUCTS_2D):
        assert L.cwa2d_struct_size(k) == C.sizeof(st)
    assert L.cwa2d_struct_size(9) == -1
    assert [L.cwa2d_kernel_name(k).decode() for k in range(7)] == [
        "hash", "scan", "scatter", "reorder", "density", "force", "wave"]
    spec = batch2d.small_batch(nscenes=2)
    bad = []
    p = spec.params.copy()
    p.h = 0.0
    bad.append((p, spec.scenes()))
    for h in (2.0 ** 20, 2.0 ** -17):   # outside the pair kernels' exact-arithmetic range
        p = spec.params.copy()
        p.h = h
        bad.append((p, spec.scenes()))
    sc = spec.scenes()
    sc[1].wave_len = 2
    bad.append((spec.params, sc))
    sc = spec.scenes()
    sc[0].n = 0
    bad.append((spec.params, sc))
    p = spec.params.copy()
    p.sph_substeps = 0
    bad.append((p, spec.scenes()))
    for prm, scn in bad:
        h = C.c_void_p()
        rc = L.cwa2d_batch_create(C.byref(prm), 2, scn, _lib.fptr(spec.xv), None, None, None,
                                  C.byref(h))
        assert rc == _lib.CWA_EINVAL and not h.value
        assert L.cwa_last_error()


def test_struct_sizes_and_version(cwa):
    from coupledwateranimation_amd import _lib

    L = cwa.lib()
    assert L.cwa_abi_version() == 1
    for k, st in enumerate(_lib.STRUCTS):
        assert L.cwa_struct_size(k) == C.sizeof(st)
    assert L.cwa_struct_size(99) == -1
    assert L.cwa_kernel_name(5).decode() == "density"


def test_reference_params(cwa):
    """Main.cpp:184-204 and the shader #defines."""
    p = cwa.reference_params()
    f = np.float32
    assert p.mass == f(0.02) and p.h == f(0.01) and p.visc == 3000 and p.rest_rho == 1000
    assert p.gas_const == 4000 and p.particle_radius == f(0.005)
    assert p.spiky_coeff == 20 and p.visc_coeff == 20
    assert list(p.gravity) == [0.0, f(-9806.65), 0.0]
    assert list(p.upper) == [f(0.48), 1.0, f(0.48)] and list(p.lower) == [0.0, f(-0.02), 0.0]
    assert (p.lambda_x, p.atten, p.beta, p.wave_type) == (f(0.01), f(0.985), f(0.001), 1.0)
    assert list(p.wave_uv_scale) == [2.0, 2.0]
    assert list(p.grid_dims) == [48, 102, 48]
    assert p.couple_band == 0.0


def test_lattice_generator(cwa):
    from coupledwateranimation_amd import configs

    a = configs.lattice(4096, (0, 0, 0), (0.3, 0.6, 0.3), 0.02, 0.5, 0.1, (2, 0, 1), 0)
    b = configs.lattice(4096, (0, 0, 0), (0.3, 0.6, 0.3), 0.02, 0.5, 0.1, (2, 0, 1), 0)
    c = configs.lattice(4096, (0, 0, 0), (0.3, 0.6, 0.3), 0.02, 0.5, 0.1, (2, 0, 1), 1)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    # site k: z fastest (15 sites), then x (15), then y from the floor
    k = np.arange(4096)
    site = np.stack([(k // 15) % 15, k // 225, k % 15], 1) * 0.02 + 0.01
    assert np.abs(a[:, :3] - site).max() <= 0.1 * 0.02 * (1 + 1e-5)
    assert np.all(a[:, 3] == 1.0)
    assert np.abs(a[:, :3] - site).max() > 0.09 * 0.02     # jitter present
    # make_cube (Main.cpp:735-759): i*s, 64 x 5 x 64, z fastest, then y, then x
    s = np.float32(2 * 0.85 * 0.005)
    r = configs.lattice(20480, (0, 0, 0), (64 * s, 5 * s, 64 * s), s, 0.0, 0.0, (2, 1, 0), 0)
    assert np.allclose(r[7, :3], [0, 0, 7 * s]) and np.allclose(r[64 * 5 * 3 + 64 * 2 + 5, :3],
                                                                 [3 * s, 2 * s, 5 * s])


def test_bumps_generator(cwa):
    from coupledwateranimation_amd import configs

    h = configs.bumps(256, 128, 8, (0.01, 0.05), (5, 40), 7)
    assert h.shape == (128, 256) and h.min() >= 0 and 0.01 <= h.max() <= 8 * 0.05
    assert np.array_equal(h, configs.bumps(256, 128, 8, (0.01, 0.05), (5, 40), 7))
    one = configs.bumps(64, 64, 1, (0.03, 0.03), (4, 4), 3)
    assert one.max() == pytest.approx(0.03, rel=0.05)


def test_workload_builders(cwa):
    from coupledwateranimation_amd import configs

    s = configs.dam_break_small()
    assert s.n == 4096 and s.wave_size == (128, 128) and s.dt == 1e-3 and s.steps == 200
    assert list(s.params.grid_dims) == [25, 25, 25]
    assert s.params.lambda_x == pytest.approx((1e-3 * 128) ** 2, rel=1e-6)
    big = configs.dam_break_large(steps=1)
    assert big.n == 1 << 20 and big.wave_size == (2048, 2048)
    assert list(big.params.grid_dims) == [113, 113, 57] and big.params.grid_ydiv == 2
    sp = big.info["spacing"]
    assert big.pos[:, 0].max() < 0.82 and big.pos[:, 1].max() < 1.0
    assert (4 / 3) * np.pi * (big.params.h / sp) ** 3 == pytest.approx(31, rel=0.02)
    assert 2 * (big.params.lambda_x + big.params.lambda_z) < 1.0   # wave CFL


@pytest.mark.parametrize("field,value", [("h", 0.0), ("cell_size", 0.001), ("rest_rho", 0.0),
                                         ("grid_dims", [0, 1, 1]), ("gas_const", 0.0)])
def test_invalid_params_rejected(cwa, field, value):
    """Validation runs before any device call, so this needs no GPU."""
    p = cwa.reference_params()
    if isinstance(value, list):
        getattr(p, field)[:] = value
    else:
        setattr(p, field, value)
    pos = np.zeros((4, 4), np.float32)
    with pytest.raises(cwa.CwaError) as e:
        cwa.Scene(p, pos)
    assert e.value.code == -1


def test_invalid_init_rejected(cwa):
    p = cwa.reference_params()
    with pytest.raises(cwa.CwaError):
        cwa.Scene(p, np.zeros((4, 4), np.float32), wave_size=(0, 8))


def test_particle_limit_rejected(cwa):
    """n above CWA_MAX_PARTICLES fails with EINVAL before any device work
    (slot indices of the neighbour records keep two flag bits)."""
    import ctypes as C

    from coupledwateranimation_amd import _lib

    p = cwa.reference_params()
    one = np.zeros(4, np.float32)
    init = _lib.Init()
    init.n = (1 << 29) + 1
    init.pos = one.ctypes.data_as(C.POINTER(C.c_float))
    init.wave_w = init.wave_h = 8
    h = C.c_void_p()
    rc = _lib.lib().cwa_scene_create(C.byref(p), C.byref(init), C.byref(h))
    assert rc == _lib.CWA_EINVAL and not h.value
    assert "CWA_MAX_PARTICLES" in _lib.lib().cwa_last_error().decode()
