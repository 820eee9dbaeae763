"""ctypes binding of the CPU oracle (oracle/_build/libcwa_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU
baseline of bench.py, never by the shipped package.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libcwa_oracle.so"

_lib = None


def build() -> Path:
    subprocess.run(["make", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run make -C oracle)")
        from coupledwateranimation_amd._lib import Params

        L = C.CDLL(str(LIB_PATH))
        P = C.POINTER(Params)
        fp = C.POINTER(C.c_float)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int32)
        sig = {
            "cwao_max_threads": (C.c_int, []),
            "cwao_exclusive_scan": (None, [ip, C.c_int, ip]),
            "cwao_blelloch_scan": (None, [ip, C.c_int]),
            "cwao_grid_build": (C.c_int, [P, fp, C.c_int, ip, ip, ip, ip]),
            "cwao_grid_build_boxes": (C.c_longlong, [fp, fp, C.c_int, fp, fp, ip, ip, ip, ip,
                                                     C.c_longlong]),
            "cwao_neighbor_counts": (C.c_int, [P, fp, C.c_int, C.c_int, ip]),
            "cwao_init_wave_f": (None, [C.c_float, fp, C.c_int, C.c_int]),
            "cwao_init_wave_d": (None, [C.c_double, dp, C.c_int, C.c_int]),
            "cwao_tex_sample_f": (C.c_float, [fp, C.c_int, C.c_int, C.c_float, C.c_float]),
            "cwao_wave_evolve_f": (None, [P, fp, fp, fp, C.c_int, C.c_int, fp, C.c_float]),
            "cwao_wave_evolve_d": (None, [P, dp, dp, dp, C.c_int, C.c_int, dp, C.c_double]),
            "cwao_scene_create_f": (C.c_void_p, [P, C.c_int, fp, fp, fp, C.c_int, C.c_int, fp,
                                                 fp, C.c_int, C.c_int]),
            "cwao_scene_create_d": (C.c_void_p, [P, C.c_int, dp, dp, dp, C.c_int, C.c_int, dp,
                                                 dp, C.c_int, C.c_int]),
            "cwao_scene_destroy": (None, [C.c_void_p]),
            "cwao_scene_set_params": (C.c_int, [C.c_void_p, P]),
            "cwao_scene_step": (C.c_int, [C.c_void_p, C.c_double, C.c_int]),
            "cwao_scene_n": (C.c_int, [C.c_void_p]),
            "cwao_scene_get_f": (C.c_int, [C.c_void_p, fp, fp, fp, fp, fp, fp, fp, fp, ip]),
            "cwao_scene_get_d": (C.c_int, [C.c_void_p, dp, dp, dp, dp, dp, dp, dp, dp, ip]),
            "cwao_scene_grid": (C.c_int, [C.c_void_p, ip, ip, ip, ip, fp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a, ctype):
    return None if a is None else a.ctypes.data_as(C.POINTER(ctype))


def exclusive_scan(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, np.int32)
    out = np.empty(x.size + 1, np.int32)
    lib().cwao_exclusive_scan(_p(x, C.c_int32), x.size, _p(out, C.c_int32))
    return out


def blelloch_scan(x: np.ndarray) -> np.ndarray:
    y = np.ascontiguousarray(x, np.int32).copy()
    lib().cwao_blelloch_scan(_p(y, C.c_int32), y.size)
    return y


def grid_build(params, pos4: np.ndarray) -> dict:
    pos4 = np.ascontiguousarray(pos4, np.float32)
    n = pos4.shape[0]
    nc = params.grid_dims[0] * params.grid_dims[1] * params.grid_dims[2]
    key = np.empty(n, np.int32)
    count = np.empty(nc, np.int32)
    offset = np.empty(nc + 1, np.int32)
    lst = np.empty(n, np.int32)
    lib().cwao_grid_build(C.byref(params), _p(pos4, C.c_float), n, _p(key, C.c_int32),
                          _p(count, C.c_int32), _p(offset, C.c_int32), _p(lst, C.c_int32))
    return {"key": key, "count": count, "offset": offset, "list": lst}


def grid_build_boxes(centre4, halfw4, grid_min, cell_size, dims) -> dict:
    """UniformGridPbdGpu3D's box grid: counts, offsets and the stable list
    (boxes ascending within each cell)."""
    x = np.ascontiguousarray(centre4, np.float32).reshape(-1, 4)
    hw = np.ascontiguousarray(halfw4, np.float32).reshape(-1, 4)
    n = x.shape[0]
    gmin = np.ascontiguousarray(grid_min, np.float32)
    cell = np.ascontiguousarray(cell_size, np.float32)
    d = np.ascontiguousarray(dims, np.int32)
    nc = int(d[0]) * int(d[1]) * int(d[2])
    count = np.empty(nc, np.int32)
    offset = np.empty(nc + 1, np.int32)
    total = lib().cwao_grid_build_boxes(_p(x, C.c_float), _p(hw, C.c_float), n,
                                        _p(gmin, C.c_float), _p(cell, C.c_float),
                                        _p(d, C.c_int32), _p(count, C.c_int32),
                                        _p(offset, C.c_int32), None, 0)
    lst = np.empty(max(int(total), 1), np.int32)
    lib().cwao_grid_build_boxes(_p(x, C.c_float), _p(hw, C.c_float), n, _p(gmin, C.c_float),
                                _p(cell, C.c_float), _p(d, C.c_int32), _p(count, C.c_int32),
                                _p(offset, C.c_int32), _p(lst, C.c_int32), int(total))
    return {"count": count, "offset": offset, "list": lst[:int(total)], "total": int(total)}


def neighbor_counts(params, pos4: np.ndarray, all_pairs: bool) -> np.ndarray:
    pos4 = np.ascontiguousarray(pos4, np.float32)
    out = np.empty(pos4.shape[0], np.int32)
    lib().cwao_neighbor_counts(C.byref(params), _p(pos4, C.c_float), pos4.shape[0],
                               1 if all_pairs else 0, _p(out, C.c_int32))
    return out


def init_wave(wave_type: float, W: int, H: int, double: bool = False) -> np.ndarray:
    if double:
        out = np.empty((H, W), np.float64)
        lib().cwao_init_wave_d(wave_type, _p(out, C.c_double), W, H)
    else:
        out = np.empty((H, W), np.float32)
        lib().cwao_init_wave_f(wave_type, _p(out, C.c_float), W, H)
    return out


def tex_sample(img: np.ndarray, u: float, v: float) -> float:
    img = np.ascontiguousarray(img, np.float32)
    H, W = img.shape
    return lib().cwao_tex_sample_f(_p(img, C.c_float), W, H, u, v)


def wave_evolve(params, h0, h1, dt=0.0, col_flux=None, double=False) -> np.ndarray:
    t, ct = (np.float64, C.c_double) if double else (np.float32, C.c_float)
    h0 = np.ascontiguousarray(h0, t)
    h1 = np.ascontiguousarray(h1, t)
    H, W = h0.shape
    out = np.empty_like(h0)
    cf = None if col_flux is None else np.ascontiguousarray(col_flux, t)
    fn = lib().cwao_wave_evolve_d if double else lib().cwao_wave_evolve_f
    fn(C.byref(params), _p(h0, ct), _p(h1, ct), _p(out, ct), W, H, _p(cf, ct), dt)
    return out


class OracleScene:
    """The oracle's coupled scene.  double=True is the FP64 twin; mode
    'grid' (27-cell search with reordering) or 'all_pairs' (the reference's
    literal O(N^2) loops)."""

    def __init__(self, params, pos, vel=None, force=None, wave_size=(64, 64), height0=None,
                 height1=None, double=False, mode="grid", threads=1):
        self.double = double
        t, ct = (np.float64, C.c_double) if double else (np.float32, C.c_float)
        self._t, self._ct = t, ct
        pos = np.ascontiguousarray(pos, t)
        n = pos.shape[0]
        W, H = wave_size
        arrs = [pos] + [None if a is None else np.ascontiguousarray(a, t)
                        for a in (vel, force, height0, height1)]
        self._keep = arrs
        create = lib().cwao_scene_create_d if double else lib().cwao_scene_create_f
        self._h = create(C.byref(params), n, _p(arrs[0], ct), _p(arrs[1], ct), _p(arrs[2], ct),
                         W, H, _p(arrs[3], ct), _p(arrs[4], ct), 0 if mode == "grid" else 1,
                         threads)
        self.n, self.W, self.H = n, W, H

    def close(self):
        if self._h:
            lib().cwao_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_params(self, params):
        lib().cwao_scene_set_params(self._h, C.byref(params))

    def step(self, dt: float, nsteps: int = 1):
        lib().cwao_scene_step(self._h, dt, nsteps)

    def state(self) -> dict:
        t, ct = self._t, self._ct
        n, W, H = self.n, self.W, self.H
        out = {k: np.empty((n, 4), t) for k in ("pos", "vel", "force")}
        for k in ("rho", "prs", "flux"):
            out[k] = np.empty(n, t)
        out["height0"] = np.empty((H, W), t)
        out["height1"] = np.empty((H, W), t)
        out["slot_of_id"] = np.empty(n, np.int32)
        get = lib().cwao_scene_get_d if self.double else lib().cwao_scene_get_f
        get(self._h, *[_p(out[k], ct) for k in ("pos", "vel", "force", "rho", "prs", "flux",
                                                "height0", "height1")],
            _p(out["slot_of_id"], C.c_int32))
        return out

    def grid(self, params) -> dict:
        n = self.n
        nc = params.grid_dims[0] * params.grid_dims[1] * params.grid_dims[2]
        ncol = params.grid_dims[0] * params.grid_dims[2]
        out = {"key": np.empty(n, np.int32), "count": np.empty(nc, np.int32),
               "offset": np.empty(nc + 1, np.int32), "list": np.empty(n, np.int32),
               "colflux": np.empty(ncol, np.float32)}
        lib().cwao_scene_grid(self._h, *[_p(out[k], C.c_int32)
                                         for k in ("key", "count", "offset", "list")],
                              _p(out["colflux"], C.c_float))
        return out
