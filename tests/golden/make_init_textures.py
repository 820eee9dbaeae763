"""Writes tests/golden/init_textures.npz: the pixels of the reference's three
wave init textures (CoupledWaterAnimation/init-textures/*.png), decoded by
coupledwateranimation_amd.files.read_png, so that GPU tests can use them
where the reference tree is absent.  Run from the repository root:

    python tests/golden/make_init_textures.py
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from coupledwateranimation_amd.files import read_png  # noqa: E402

SRC = Path("/root/reference/CoupledWaterAnimation/init-textures")
NAMES = ("breaking-wave", "splash", "wake")


def main() -> None:
    arrays = {name.replace("-", "_"): read_png(SRC / f"{name}.png") for name in NAMES}
    out = Path(__file__).with_name("init_textures.npz")
    np.savez_compressed(out, **arrays)
    for k, v in arrays.items():
        print(k, v.shape, v.dtype, int(v[..., 0].min()), int(v[..., 0].max()))
    print("wrote", out)


if __name__ == "__main__":
    main()
