"""Python mirror of the reference's step API over the C ABI.

``Scene`` plays the role the reference gives to its particle SSBO plus the
StencilImage2DTripleBuffered Module (CoupledWaterAnimation/Main.cpp:764-794,
StencilImage2DTripleBuffered.h:8-41):

    Scene(params, pos, ...)      init_particles() + Module::sInitAll()
    scene.reinit(...)            the 'R' reset, Main.cpp:677-684
    scene.step(dt, n)            idle(), Main.cpp:530-562
    scene.wave_reinit()          StencilImage2DTripleBuffered::Reinit
    scene.wave_compute(dt)       StencilImage2DTripleBuffered::Compute
    scene.height(age)            GetReadImage(age) / GetWriteImage()
    scene.particles()            the particle SSBO handle (device pointers)

Errors raise ``CwaError`` with the library's message.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


def reference_params() -> L.Params:
    """cwa_params_reference(): the reference's own constants."""
    p = L.Params()
    L.lib().cwa_params_reference(C.byref(p))
    return p


def set_lane_split_limit(max_n: int) -> int:
    """cwa_set_lane_split_limit: scenes of up to max_n particles run the
    density pass (and the pair adjoint of tapes created afterwards) with four
    lanes per particle, larger ones with one; max_n < 0 restores the default
    (2^18).  Returns the previous limit."""
    prev = C.c_longlong()
    L.check(L.lib().cwa_set_lane_split_limit(int(max_n), C.byref(prev)))
    return int(prev.value)


def _f32(a, shape=None):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None and a.size != int(np.prod(shape)):
        raise ValueError(f"expected {shape} elements, got {a.shape}")
    return a


def _hip():
    return _HipRuntime.get()


class _HipRuntime:
    """hipMemcpy from the HIP runtime, for copying views to the host."""

    _inst = None

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so")
        self.rt.hipMemcpy.restype = C.c_int
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]

    @classmethod
    def get(cls):
        if cls._inst is None:
            cls._inst = cls()
        return cls._inst

    def d2h(self, ptr: int, nbytes: int, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        if nbytes:
            rc = self.rt.hipMemcpy(out.ctypes.data, ptr, nbytes, 2)
            if rc != 0:
                raise RuntimeError(f"hipMemcpy failed ({rc})")
        return out


class Scene:
    def __init__(self, params: L.Params, pos, vel=None, force=None, wave_size=(64, 64),
                 height0=None, height1=None):
        self._lib = L.lib()
        self._h = C.c_void_p()
        self._keep = None
        init = self._make_init(pos, vel, force, wave_size, height0, height1)
        L.check(self._lib.cwa_scene_create(C.byref(params), C.byref(init), C.byref(self._h)))

    def _make_init(self, pos, vel, force, wave_size, height0, height1) -> L.Init:
        pos = _f32(pos)
        n = 0 if pos is None else pos.size // 4
        W, H = int(wave_size[0]), int(wave_size[1])
        vel = _f32(vel, (n, 4))
        force = _f32(force, (n, 4))
        height0 = _f32(height0, (H, W))
        height1 = _f32(height1, (H, W))
        self._keep = (pos, vel, force, height0, height1)
        init = L.Init()
        init.n = n
        init.pos = L.fptr(pos)
        init.vel = L.fptr(vel)
        init.force = L.fptr(force)
        init.wave_w, init.wave_h = W, H
        init.height0 = L.fptr(height0)
        init.height1 = L.fptr(height1)
        return init

    # ---- lifetime -------------------------------------------------------
    def close(self) -> None:
        if self._h:
            L.check(self._lib.cwa_scene_destroy(self._h))
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reinit(self, pos, vel=None, force=None, wave_size=(64, 64), height0=None,
               height1=None) -> None:
        init = self._make_init(pos, vel, force, wave_size, height0, height1)
        L.check(self._lib.cwa_scene_reinit(self._h, C.byref(init)))

    # ---- parameters -----------------------------------------------------
    def params(self) -> L.Params:
        p = L.Params()
        L.check(self._lib.cwa_scene_get_params(self._h, C.byref(p)))
        return p

    def set_params(self, p: L.Params) -> None:
        L.check(self._lib.cwa_scene_set_params(self._h, C.byref(p)))

    # ---- stepping -------------------------------------------------------
    def step(self, dt: float, nsteps: int = 1, graph: bool = False) -> None:
        fn = self._lib.cwa_scene_step_graph if graph else self._lib.cwa_scene_step
        L.check(fn(self._h, float(dt), int(nsteps)))

    def step_profiled(self, dt: float, nsteps: int) -> dict:
        kt = L.KernelTimes()
        L.check(self._lib.cwa_scene_step_profiled(self._h, float(dt), int(nsteps), C.byref(kt)))
        out = {}
        for k in range(L.CWA_NUM_KERNELS):
            if kt.launches[k]:
                out[self._lib.cwa_kernel_name(k).decode()] = (kt.ms[k], kt.launches[k])
        return out

    def sync(self) -> None:
        L.check(self._lib.cwa_scene_sync(self._h))

    def stream(self) -> int:
        return self._lib.cwa_scene_stream(self._h) or 0

    # ---- wave Module ----------------------------------------------------
    def wave_reinit(self) -> None:
        L.check(self._lib.cwa_scene_wave_reinit(self._h))

    def wave_reinit_from_texture(self, tex) -> None:
        tex = _f32(tex)
        th, tw = tex.shape
        L.check(self._lib.cwa_scene_wave_reinit_from_texture(self._h, L.fptr(tex), tw, th))

    def wave_compute(self, dt: float) -> None:
        L.check(self._lib.cwa_scene_wave_compute(self._h, float(dt)))

    def height(self, age: int = 0):
        """(device pointer, width, height, pitch) of GetReadImage(age)."""
        p = C.c_void_p()
        w, h, pitch = C.c_int(), C.c_int(), C.c_int()
        L.check(self._lib.cwa_scene_height(self._h, age, C.byref(p), C.byref(w), C.byref(h),
                                           C.byref(pitch)))
        return p.value, w.value, h.value, pitch.value

    # ---- handles --------------------------------------------------------
    def particles(self) -> L.ParticleView:
        v = L.ParticleView()
        L.check(self._lib.cwa_scene_particles(self._h, C.byref(v)))
        return v

    def prepare_graph(self, dt: float) -> bool:
        """cwa_scene_prepare_graph: capture the step graph for dt without
        running a step; False when none can be captured yet."""
        c = C.c_int()
        L.check(self._lib.cwa_scene_prepare_graph(self._h, float(dt), C.byref(c)))
        return bool(c.value)

    def particles_written(self) -> None:
        """cwa_scene_particles_written: particle state was written through
        the device pointers of particles(); the next step hashes afresh."""
        L.check(self._lib.cwa_scene_particles_written(self._h))

    def grid(self) -> L.GridView:
        v = L.GridView()
        L.check(self._lib.cwa_scene_grid(self._h, C.byref(v)))
        return v

    def grid_host(self) -> dict:
        """Grid arrays of the last step copied to host memory."""
        self.sync()
        g = self.grid()
        n, nc = g.n, g.num_cells
        hip = _hip()
        ncol = g.dims[0] * g.dims[2]
        return {
            "dims": tuple(g.dims), "key": hip.d2h(g.key, 4 * n, np.int32, n),
            "count": hip.d2h(g.count, 4 * nc, np.int32, nc),
            "offset": hip.d2h(g.offset, 4 * (nc + 1), np.int32, nc + 1),
            "list": hip.d2h(g.list, 4 * n, np.int32, n),
            "colflux": hip.d2h(g.flux, 4 * ncol, np.float32, ncol),
        }

    def slot_ids(self) -> np.ndarray:
        """Original particle index held by each (sorted) slot."""
        self.sync()
        v = self.particles()
        return _hip().d2h(v.id, 4 * v.n, np.int32, v.n)

    def neighbour_counts(self) -> np.ndarray:
        """Neighbour count (self included) of every particle, original order,
        as the last step's density pass found it."""
        n = self.particles().n
        out = np.zeros(n, np.int32)
        if n:
            L.check(self._lib.cwa_scene_neighbour_counts(self._h,
                                                         out.ctypes.data_as(C.POINTER(C.c_int32))))
        return out

    def readback(self) -> dict:
        """State in original particle order plus the two newest heights."""
        v = self.particles()
        n = v.n
        _, W, H, _ = self.height(0)
        out = {
            "pos": np.empty((n, 4), np.float32), "vel": np.empty((n, 4), np.float32),
            "force": np.empty((n, 4), np.float32), "rho": np.empty(n, np.float32),
            "prs": np.empty(n, np.float32), "height0": np.empty((H, W), np.float32),
            "height1": np.empty((H, W), np.float32),
        }
        hs = L.HostState()
        for k in ("pos", "vel", "force", "rho", "prs", "height0", "height1"):
            setattr(hs, k, L.fptr(out[k]))
        L.check(self._lib.cwa_scene_readback(self._h, C.byref(hs)))
        return out
