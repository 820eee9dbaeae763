"""Initial-condition and scene files.

* PNG init textures: the reference loads `init-textures/*.png` through
  FreeImage as 8-bit RGBA with the bottom image row first
  (CoupledWaterAnimation/LoadTexture.cpp:5-33, `FreeImage_ConvertToRawBits`
  with top-down FALSE, uploaded as GL_RGBA8), and `InitFromImage` stores
  the red channel of texel (2x, y) as the height (wave_comp.glsl:76-80).
  `init_texture` returns that red channel as float32 in GL row order, ready
  for `Scene.wave_reinit_from_texture`.  The PNG reader is self-contained
  (zlib + the five scanline filters, chunk CRCs checked): 8-bit greyscale,
  grey+alpha, RGB and RGBA, non-interlaced.

* Scene files: `save_scene` / `load_scene` keep a SceneSpec (parameters,
  particle arrays, heightfields, time step) in one .npz written with
  numpy's own format; loading uses `allow_pickle=False`, so a file runs no
  code.  The parameters travel as JSON text.
"""
from __future__ import annotations

import json
import struct
import zlib
from pathlib import Path

import numpy as np

from . import _lib as L

PNG_SIGNATURE = b"\x89PNG\r\n\x1a\n"
_CHANNELS = {0: 1, 2: 3, 4: 2, 6: 4}


def _paeth(a: int, b: int, c: int) -> int:
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    return b if pb <= pc else c


def unfilter(raw: bytes, width: int, height: int, bpp: int) -> np.ndarray:
    """Undo the per-scanline filters (PNG specification 9.2-9.4)."""
    stride = width * bpp
    out = np.zeros((height, stride), np.uint8)
    prev = np.zeros(stride, np.int32)
    pos = 0
    for y in range(height):
        ftype = raw[pos]
        line = np.frombuffer(raw, np.uint8, stride, pos + 1).astype(np.int32)
        pos += 1 + stride
        cur = np.zeros(stride, np.int32)
        if ftype == 0:
            cur[:] = line
        elif ftype == 2:
            cur[:] = (line + prev) & 0xFF
        else:
            for x in range(stride):
                a = cur[x - bpp] if x >= bpp else 0
                b = prev[x]
                c = prev[x - bpp] if x >= bpp else 0
                if ftype == 1:
                    pred = a
                elif ftype == 3:
                    pred = (a + b) >> 1
                elif ftype == 4:
                    pred = _paeth(a, b, c)
                else:
                    raise ValueError(f"unknown PNG filter type {ftype}")
                cur[x] = (line[x] + pred) & 0xFF
        out[y] = cur
        prev = cur
    return out


def read_png(path) -> np.ndarray:
    """(H, W, C) uint8 pixels, the first row the top of the image."""
    data = Path(path).read_bytes()
    if data[:8] != PNG_SIGNATURE:
        raise ValueError(f"{path}: not a PNG file")
    pos, ihdr, idat = 8, None, []
    while pos < len(data):
        (n,) = struct.unpack(">I", data[pos:pos + 4])
        ctype = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        if zlib.crc32(ctype + body) & 0xFFFFFFFF != crc:
            raise ValueError(f"{path}: CRC mismatch in chunk {ctype!r}")
        if ctype == b"IHDR":
            ihdr = struct.unpack(">IIBBBBB", body)
        elif ctype == b"IDAT":
            idat.append(body)
        elif ctype == b"IEND":
            break
        pos += 12 + n
    if ihdr is None:
        raise ValueError(f"{path}: no IHDR chunk")
    w, h, depth, color, comp, filt, interlace = ihdr
    if depth != 8 or color not in _CHANNELS or comp != 0 or filt != 0 or interlace != 0:
        raise ValueError(f"{path}: unsupported PNG (depth {depth}, colour type {color}, "
                         f"interlace {interlace})")
    ch = _CHANNELS[color]
    pix = unfilter(zlib.decompress(b"".join(idat)), w, h, ch)
    return pix.reshape(h, w, ch)


def write_png(path, pixels: np.ndarray, filters=(0,)) -> None:
    """Write (H, W, C) uint8 pixels (C = 1, 2, 3 or 4); scanline y uses
    filters[y % len(filters)].  Used to test the reader."""
    pix = np.ascontiguousarray(pixels, np.uint8)
    h, w, ch = pix.shape
    color = {1: 0, 2: 4, 3: 2, 4: 6}[ch]
    bpp, stride = ch, w * ch
    rows = pix.reshape(h, stride).astype(np.int32)
    raw = bytearray()
    prev = np.zeros(stride, np.int32)
    for y in range(h):
        f = filters[y % len(filters)]
        cur = rows[y]
        a = np.concatenate([np.zeros(bpp, np.int32), cur[:-bpp]])
        c = np.concatenate([np.zeros(bpp, np.int32), prev[:-bpp]])
        if f == 0:
            line = cur
        elif f == 1:
            line = cur - a
        elif f == 2:
            line = cur - prev
        elif f == 3:
            line = cur - ((a + prev) >> 1)
        else:
            line = cur - np.array([_paeth(int(x), int(y_), int(z)) for x, y_, z in
                                   zip(a, prev, c)], np.int32)
        raw.append(f)
        raw += (line & 0xFF).astype(np.uint8).tobytes()
        prev = cur

    def chunk(t, body):
        return (struct.pack(">I", len(body)) + t + body +
                struct.pack(">I", zlib.crc32(t + body) & 0xFFFFFFFF))

    ihdr = struct.pack(">IIBBBBB", w, h, 8, color, 0, 0, 0)
    Path(path).write_bytes(PNG_SIGNATURE + chunk(b"IHDR", ihdr) +
                           chunk(b"IDAT", zlib.compress(bytes(raw), 9)) + chunk(b"IEND", b""))


def texture_from_pixels(pix: np.ndarray) -> np.ndarray:
    """The red channel as the reference's GL texture holds it: /255, the
    bottom image row first (texture row 0)."""
    red = pix[..., 0].astype(np.float32) / np.float32(255.0)
    return np.ascontiguousarray(red[::-1])


def init_texture(path) -> np.ndarray:
    """(H, W) float32 init texture from a PNG, for wave_reinit_from_texture."""
    return texture_from_pixels(read_png(path))


# ---------------------------------------------------------------------------
# scene files
# ---------------------------------------------------------------------------
def _params_to_json(p) -> str:
    out = {}
    for name, _ in type(p)._fields_:
        v = getattr(p, name)
        out[name] = list(v) if hasattr(v, "__len__") else v
    return json.dumps({"type": type(p).__name__, "fields": out})


def _params_from_json(text: str):
    d = json.loads(text)
    cls = {"Params": L.Params, "Params2D": L.Params2D}[d["type"]]
    p = cls()
    for name, v in d["fields"].items():
        if isinstance(v, list):
            arr = getattr(p, name)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(p, name, v)
    return p


def save_scene(path, spec, state: dict | None = None) -> None:
    """A SceneSpec (and optionally a Scene.readback() state that replaces its
    particle arrays and heightfields) as .npz."""
    pos, vel = spec.pos, spec.vel
    h0, h1 = spec.height0, spec.height1
    if state is not None:
        pos, vel = state["pos"], state["vel"]
        h0, h1 = state["height0"], state["height1"]
    meta = {"name": spec.name, "wave_size": list(spec.wave_size), "dt": spec.dt,
            "steps": spec.steps, "info": spec.info}
    np.savez(path, pos=np.asarray(pos, np.float32), vel=np.asarray(vel, np.float32),
             height0=np.asarray(h0, np.float32), height1=np.asarray(h1, np.float32),
             params=np.frombuffer(_params_to_json(spec.params).encode(), np.uint8),
             meta=np.frombuffer(json.dumps(meta).encode(), np.uint8))


def load_scene(path):
    """The SceneSpec stored by save_scene (no code runs while loading)."""
    from .configs import SceneSpec

    with np.load(path, allow_pickle=False) as z:
        meta = json.loads(z["meta"].tobytes().decode())
        params = _params_from_json(z["params"].tobytes().decode())
        return SceneSpec(meta["name"], params, z["pos"].copy(), z["vel"].copy(),
                         tuple(meta["wave_size"]), z["height0"].copy(), z["height1"].copy(),
                         meta["dt"], meta["steps"], meta["info"])
