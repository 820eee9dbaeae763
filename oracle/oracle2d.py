"""ctypes binding of the batched 2D oracle (oracle/_build/libcwa_oracle2d.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU
baseline of bench.py, never by the shipped package.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libcwa_oracle2d.so"

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run make -C oracle)")
        from coupledwateranimation_amd._lib import Params2D, Scene2D

        L = C.CDLL(str(LIB_PATH))
        P = C.POINTER(Params2D)
        fp = C.POINTER(C.c_float)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int32)
        sig = {
            "cwao2_create": (C.c_void_p, [P, C.c_int, C.POINTER(Scene2D), fp, ip, fp, fp,
                                          C.c_int]),
            "cwao2_destroy": (None, [C.c_void_p]),
            "cwao2_set_params": (C.c_int, [C.c_void_p, P]),
            "cwao2_step": (C.c_int, [C.c_void_p, C.c_double, C.c_int, C.c_int]),
            "cwao2_get": (C.c_int, [C.c_void_p, C.c_int, dp, dp, ip, dp, dp]),
            "cwao2_grid": (C.c_int, [C.c_void_p, C.c_int, ip, ip]),
            "cwao2_colsum": (C.c_int, [C.c_void_p, C.c_int, dp]),
            "cwao2_dims": (None, [P, ip, ip]),
            "cwao2_W_d": (C.c_double, [P, C.c_double]),
            "cwao2_Wgrad_d": (C.c_double, [P, C.c_double]),
            "cwao2_sdf_d": (C.c_double, [P, C.c_double, C.c_double, dp]),
            "cwao2_wave_at_d": (C.c_double, [P, dp, C.c_int, C.c_double]),
            "cwao2_column_sum_f": (C.c_float, [fp, C.c_int]),
            "cwao2_shallow_d": (None, [dp, dp, C.c_int, C.c_double, C.c_double, C.c_int]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a, ctype):
    return None if a is None else a.ctypes.data_as(C.POINTER(ctype))


def dims(params) -> tuple[int, int]:
    nx, ny = C.c_int32(), C.c_int32()
    lib().cwao2_dims(C.byref(params), C.byref(nx), C.byref(ny))
    return nx.value, ny.value


def W(params, r: float) -> float:
    return lib().cwao2_W_d(C.byref(params), float(r))


def Wgrad(params, r: float) -> float:
    return lib().cwao2_Wgrad_d(C.byref(params), float(r))


def sdf(params, x: float, y: float):
    n = np.zeros(2, np.float64)
    d = lib().cwao2_sdf_d(C.byref(params), float(x), float(y), _p(n, C.c_double))
    return d, n


def wave_at(params, w, x: float) -> float:
    w = np.ascontiguousarray(w, np.float64)
    return lib().cwao2_wave_at_d(C.byref(params), _p(w, C.c_double), int(w.size), float(x))


def column_sum(f) -> float:
    f = np.ascontiguousarray(f, np.float32)
    return lib().cwao2_column_sum_f(_p(f, C.c_float), int(f.size))


def shallow(h, uh, lam: float, dx: float, nsub: int):
    """nsub Shallow1D substeps of (h, uh) in double."""
    h = np.array(h, np.float64)
    uh = np.array(uh, np.float64)
    lib().cwao2_shallow_d(_p(h, C.c_double), _p(uh, C.c_double), int(h.size), float(lam),
                          float(dx), int(nsub))
    return h, uh


class OracleBatch:
    """A batch on the CPU: float (the GPU's arithmetic) or double."""

    def __init__(self, spec, double: bool = False, threads: int = 8):
        self.spec = spec
        self.threads = threads
        xv = np.ascontiguousarray(spec.xv, np.float32)
        en = None if spec.enabled is None else np.ascontiguousarray(spec.enabled, np.int32)
        h0 = np.ascontiguousarray(spec.h0, np.float32)
        h1 = np.ascontiguousarray(spec.h1, np.float32)
        self._scenes = spec.scenes()
        self._h = lib().cwao2_create(C.byref(spec.params), spec.nscenes, self._scenes,
                                     _p(xv, C.c_float), _p(en, C.c_int32), _p(h0, C.c_float),
                                     _p(h1, C.c_float), 1 if double else 0)
        if not self._h:
            raise ValueError("oracle rejected the batch")

    def close(self):
        if self._h:
            lib().cwao2_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_params(self, params):
        lib().cwao2_set_params(self._h, C.byref(params))

    def step(self, dt: float, nsteps: int = 1):
        lib().cwao2_step(self._h, float(dt), int(nsteps), int(self.threads))

    def scene_slots(self, s: int) -> dict:
        """Scene s in slot order (cell-sorted, as the device keeps it)."""
        n, Lw = int(self.spec.n[s]), int(self.spec.wave_len[s])
        xv = np.zeros((n, 4))
        rp = np.zeros((n, 2))
        idf = np.zeros(n, np.int32)
        h0 = np.zeros(Lw)
        h1 = np.zeros(Lw)
        rc = lib().cwao2_get(self._h, s, _p(xv, C.c_double), _p(rp, C.c_double),
                             _p(idf, C.c_int32), _p(h0, C.c_double), _p(h1, C.c_double))
        assert rc == 0
        return {"xv": xv, "rp": rp, "idf": idf, "h0": h0, "h1": h1}

    def slots(self) -> dict:
        """All scenes concatenated in slot order."""
        parts = [self.scene_slots(s) for s in range(self.spec.nscenes)]
        return {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}

    def grid(self, s: int) -> dict:
        nx, ny = dims(self.spec.params)
        n = int(self.spec.n[s])
        key = np.zeros(n, np.int32)
        off = np.zeros(nx * ny + 1, np.int32)
        assert lib().cwao2_grid(self._h, s, _p(key, C.c_int32), _p(off, C.c_int32)) == 0
        return {"key": key, "offset": off}

    def state(self) -> dict:
        """Original order (scene-major, id within scene), as Batch2D.readback."""
        sl = self.slots()
        pbeg = np.repeat(self.spec.pbeg(), self.spec.n)
        gid = pbeg + (sl["idf"] & 0x3FFFFFFF)
        out = {"xv": np.zeros_like(sl["xv"]), "rp": np.zeros_like(sl["rp"]),
               "enabled": np.zeros(sl["idf"].size, np.int32), "h0": sl["h0"], "h1": sl["h1"]}
        out["xv"][gid] = sl["xv"]
        out["rp"][gid] = sl["rp"]
        out["enabled"][gid] = (sl["idf"] >> 30) & 1
        return out
