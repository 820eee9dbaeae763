"""Initial-condition and scene files (coupledwateranimation_amd/files.py).

PNG reader: pinned by round trips through an independent writer for every
scanline filter (None, Sub, Up, Average, Paeth) and colour type, with chunk
CRCs checked on read, and by the reference's own init textures decoding to
the committed pixels (tests/golden/init_textures.npz, written by
tests/golden/make_init_textures.py).  Texture orientation follows the
reference's loader (bottom row first, red / 255).  Scene files: round trip,
and loading never unpickles.
"""
from pathlib import Path

import numpy as np
import pytest

from coupledwateranimation_amd import files

REF_TEX = Path("/root/reference/CoupledWaterAnimation/init-textures")
GOLD = Path(__file__).with_name("golden") / "init_textures.npz"


@pytest.mark.parametrize("channels", [1, 2, 3, 4])
@pytest.mark.parametrize("filters", [(0,), (1,), (2,), (3,), (4,), (0, 1, 2, 3, 4)])
def test_png_round_trip(tmp_path, channels, filters):
    rng = np.random.default_rng(channels * 10 + len(filters))
    pix = rng.integers(0, 256, (13, 17, channels), dtype=np.uint8)
    f = tmp_path / "t.png"
    files.write_png(f, pix, filters)
    assert np.array_equal(files.read_png(f), pix)


def test_png_crc_checked(tmp_path):
    pix = np.full((4, 4, 3), 7, np.uint8)
    f = tmp_path / "t.png"
    files.write_png(f, pix)
    b = bytearray(f.read_bytes())
    b[40] ^= 0xFF                      # inside the IDAT payload
    f.write_bytes(bytes(b))
    with pytest.raises(ValueError):
        files.read_png(f)


@pytest.mark.skipif(not REF_TEX.exists(), reason="reference tree not present")
def test_reference_init_textures_decode_to_fixture():
    gold = np.load(GOLD, allow_pickle=False)
    for name in ("breaking-wave", "splash", "wake"):
        pix = files.read_png(REF_TEX / f"{name}.png")
        assert np.array_equal(pix, gold[name.replace("-", "_")])


def test_texture_orientation_and_scale():
    pix = np.zeros((3, 2, 3), np.uint8)
    pix[0, :, 0] = 255      # top image row
    pix[2, 1, 0] = 51       # bottom row, second pixel
    tex = files.texture_from_pixels(pix)
    assert tex.dtype == np.float32 and tex.shape == (3, 2)
    assert np.all(tex[2] == 1.0)                   # top row is the last texture row
    assert tex[0, 1] == np.float32(51) / np.float32(255)


def test_scene_file_round_trip(tmp_path):
    from coupledwateranimation_amd import configs

    spec = configs.dam_break_small()
    f = tmp_path / "scene.npz"
    files.save_scene(f, spec)
    back = files.load_scene(f)
    assert back.name == spec.name and back.wave_size == spec.wave_size
    assert back.dt == spec.dt and back.steps == spec.steps
    for k in ("pos", "vel", "height0", "height1"):
        assert np.array_equal(getattr(back, k), getattr(spec, k))
    for name, _ in type(spec.params)._fields_:
        a, b = getattr(back.params, name), getattr(spec.params, name)
        assert (list(a) == list(b)) if hasattr(a, "__len__") else a == b, name


def test_scene_file_never_unpickles(tmp_path):
    f = tmp_path / "bad.npz"
    np.savez(f, pos=np.array([object()], dtype=object))
    with pytest.raises(Exception):
        files.load_scene(f)
