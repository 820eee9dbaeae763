"""Batched 2D coupled step: many independent SPH + 1D-wave scenes advanced
together (include/cwa2d.h, csrc/cwa_batch2d.hip).

Per scene this is the SphWave2D application's step: SphUgrid::Compute
(SphWave2D/StencilBuffer.cpp:150-179: grid, density, forces per SPH
substep) then ImageStencil::Compute (SphWave2D/StencilImage2D.cpp:141-164:
Wave1D substeps).  ``Batch2D`` mirrors those modules' use:

    spec = batch2d.batch_c3()                 # 256 scenes, 1k..64k particles
    with spec.make() as b:                    # Init / Reinit
        b.tune(spec.dt)                       # autotuner (launch shapes)
        b.step(spec.dt, 500, graph=True)      # Compute, replayed from a hipGraph
        st = b.readback()                     # GetReadBuffer / GetReadImage
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .scene import _hip

PI_SHADER = 3.14159265  # M_PI as written in SphWaveKoschier2D_grid_cs.glsl:131


def kernel_names() -> list[str]:
    lib = L.lib()
    return [lib.cwa2d_kernel_name(k).decode() for k in range(L.CWA2D_NUM_KERNELS)]


def reference_params() -> L.Params2D:
    """The SphWave2D application's constants (cwa2d_params_reference)."""
    p = L.Params2D()
    L.check(L.lib().cwa2d_params_reference(C.byref(p)))
    return p


def wave_init_reference(n: int) -> tuple[np.ndarray, np.ndarray]:
    """Wave1D InitWave modes 0 and 1 (Wave1D_cs.glsl:96-118): (h0, h1)."""
    h0 = np.zeros(n, np.float32)
    h1 = np.zeros(n, np.float32)
    L.check(L.lib().cwa2d_wave_init_reference(int(n), L.fptr(h0), L.fptr(h1)))
    return h0, h1


def wave_init_shallow(n: int, view_height: float = 9.6) -> tuple[np.ndarray, np.ndarray]:
    """Shallow1D InitWave (Shallow1D_cs.glsl:117-131): (h, uh)."""
    h = np.zeros(n, np.float32)
    uh = np.zeros(n, np.float32)
    L.check(L.lib().cwa2d_wave_init_shallow(int(n), float(view_height), L.fptr(h), L.fptr(uh)))
    return h, uh


def psi_for(h: float, rest_rho: float) -> float:
    """PSI = REST_DENS / (1.5 k2), k2 = 40/(7 pi H^2), in float
    (SphWaveKoschier2D_grid_cs.glsl:135-136)."""
    f = np.float32
    k2 = f(40.0) / (f(7.0) * f(PI_SHADER) * (f(h) * f(h)))
    return float(f(rest_rho) / (f(1.5) * k2))


@dataclass
class BatchSpec:
    """A batch: parameters, per-scene sizes, initial state, time step."""

    params: L.Params2D
    n: np.ndarray                 # particles per scene
    wave_len: np.ndarray          # wave cells per scene
    wave_substeps: np.ndarray
    mass: np.ndarray
    lam: np.ndarray
    xv: np.ndarray                # (sum n, 4) float32: x, y, vx, vy
    h0: np.ndarray                # (sum L,) float32
    h1: np.ndarray
    dt: float
    steps: int
    enabled: np.ndarray | None = None
    name: str = ""
    extra: dict = field(default_factory=dict)
    wave_model: np.ndarray | None = None   # per scene; None: all leapfrog
    wave_dx: np.ndarray | None = None      # per scene; shallow-water cell size

    @property
    def nscenes(self) -> int:
        return int(self.n.size)

    @property
    def n_total(self) -> int:
        return int(self.n.sum())

    @property
    def wave_total(self) -> int:
        return int(self.wave_len.sum())

    def scenes(self):
        arr = (L.Scene2D * self.nscenes)()
        for s in range(self.nscenes):
            arr[s].n = int(self.n[s])
            arr[s].wave_len = int(self.wave_len[s])
            arr[s].wave_substeps = int(self.wave_substeps[s])
            arr[s].mass = float(self.mass[s])
            arr[s].lambda_ = float(self.lam[s])
            arr[s].wave_model = 0 if self.wave_model is None else int(self.wave_model[s])
            arr[s].wave_dx = 1.0 if self.wave_dx is None else float(self.wave_dx[s])
        return arr

    def pbeg(self) -> np.ndarray:
        return np.concatenate([[0], np.cumsum(self.n)[:-1]]).astype(np.int64)

    def wbeg(self) -> np.ndarray:
        return np.concatenate([[0], np.cumsum(self.wave_len)[:-1]]).astype(np.int64)

    def make(self) -> "Batch2D":
        return Batch2D(self.params, self.scenes(), self.xv, self.enabled, self.h0, self.h1)


class Batch2D:
    """The device batch (cwa2d_batch)."""

    def __init__(self, params: L.Params2D, scenes, xv, enabled=None, h0=None, h1=None):
        self._lib = L.lib()
        self._h = C.c_void_p()
        self._ns = len(scenes)
        self._n = np.array([scenes[s].n for s in range(self._ns)], np.int64)
        self._L = np.array([scenes[s].wave_len for s in range(self._ns)], np.int64)
        xv = np.ascontiguousarray(xv, np.float32).reshape(-1, 4)
        if xv.shape[0] != int(self._n.sum()):
            raise ValueError("xv rows must equal the total particle count")
        en = None if enabled is None else np.ascontiguousarray(enabled, np.int32)
        h0 = None if h0 is None else np.ascontiguousarray(h0, np.float32)
        h1 = None if h1 is None else np.ascontiguousarray(h1, np.float32)
        L.check(self._lib.cwa2d_batch_create(C.byref(params), self._ns, scenes, L.fptr(xv),
                                             L.iptr(en), L.fptr(h0), L.fptr(h1),
                                             C.byref(self._h)))

    def close(self) -> None:
        if self._h:
            L.check(self._lib.cwa2d_batch_destroy(self._h))
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

    @property
    def n_total(self) -> int:
        return int(self._n.sum())

    def reinit(self, xv, enabled=None, h0=None, h1=None) -> None:
        xv = np.ascontiguousarray(xv, np.float32)
        en = None if enabled is None else np.ascontiguousarray(enabled, np.int32)
        h0 = None if h0 is None else np.ascontiguousarray(h0, np.float32)
        h1 = None if h1 is None else np.ascontiguousarray(h1, np.float32)
        L.check(self._lib.cwa2d_batch_reinit(self._h, L.fptr(xv), L.iptr(en), L.fptr(h0),
                                             L.fptr(h1)))

    def params(self) -> L.Params2D:
        p = L.Params2D()
        L.check(self._lib.cwa2d_batch_get_params(self._h, C.byref(p)))
        return p

    def set_params(self, p: L.Params2D) -> None:
        L.check(self._lib.cwa2d_batch_set_params(self._h, C.byref(p)))

    def step(self, dt: float, nsteps: int = 1, graph: bool = False) -> None:
        fn = self._lib.cwa2d_batch_step_graph if graph else self._lib.cwa2d_batch_step
        L.check(fn(self._h, float(dt), int(nsteps)))

    def step_profiled(self, dt: float) -> dict:
        t = L.KernelTimes2D()
        L.check(self._lib.cwa2d_batch_step_profiled(self._h, float(dt), C.byref(t)))
        return dict(zip(kernel_names(), list(t.ms)))

    def sync(self) -> None:
        L.check(self._lib.cwa2d_batch_sync(self._h))

    def stream(self) -> int:
        s = C.c_void_p()
        L.check(self._lib.cwa2d_batch_stream(self._h, C.byref(s)))
        return s.value or 0

    def tuning(self) -> dict:
        t = L.Tuning2D()
        L.check(self._lib.cwa2d_batch_get_tuning(self._h, C.byref(t)))
        return t.as_dict()

    def set_tuning(self, **kw) -> None:
        t = L.Tuning2D()
        L.check(self._lib.cwa2d_batch_get_tuning(self._h, C.byref(t)))
        for k, v in kw.items():
            setattr(t, k, v)
        L.check(self._lib.cwa2d_batch_set_tuning(self._h, C.byref(t)))

    def tune(self, dt: float, steps: int = 3) -> dict:
        t = L.Tuning2D()
        L.check(self._lib.cwa2d_batch_tune(self._h, float(dt), int(steps), C.byref(t)))
        return t.as_dict()

    def splash(self, scene: int = -1, view_height: float = 9.6) -> None:
        """The reference's Splash button on the shallow-water scenes."""
        L.check(self._lib.cwa2d_batch_splash(self._h, int(scene), float(view_height)))

    def render(self, points_xy_ptr: int | None, wave_ptr: int | None) -> None:
        """Copy (x, y) per slot and the newest wave level into device memory
        owned by the caller (e.g. torch tensors' data_ptr())."""
        L.check(self._lib.cwa2d_batch_render(self._h, points_xy_ptr or None, wave_ptr or None))

    def views(self) -> L.Views2D:
        v = L.Views2D()
        L.check(self._lib.cwa2d_batch_views(self._h, C.byref(v)))
        return v

    def slots(self) -> dict:
        """Device state in slot order (cell-sorted within each scene)."""
        self.sync()
        v = self.views()
        hip = _hip()
        n, nc = v.n_total, v.cells_total
        return {
            "xv": hip.d2h(v.xv, 16 * n, np.float32, 4 * n).reshape(n, 4),
            "rp": hip.d2h(v.rp, 8 * n, np.float32, 2 * n).reshape(n, 2),
            "idf": hip.d2h(v.idf, 4 * n, np.int32, n),
            "key": hip.d2h(v.key, 4 * n, np.int32, n),
            "offset": hip.d2h(v.offset, 4 * (nc + 1), np.int32, nc + 1),
            "h0": hip.d2h(v.h0, 4 * v.wave_total, np.float32, v.wave_total),
            "h1": hip.d2h(v.h1, 4 * v.wave_total, np.float32, v.wave_total),
            "nx": v.grid_nx, "ny": v.grid_ny,
        }

    def readback(self) -> dict:
        """State in original order (scene-major, particle id within scene)."""
        n, w = self.n_total, int(self._L.sum())
        xv = np.zeros((n, 4), np.float32)
        rp = np.zeros((n, 2), np.float32)
        en = np.zeros(n, np.int32)
        h0 = np.zeros(w, np.float32)
        h1 = np.zeros(w, np.float32)
        L.check(self._lib.cwa2d_batch_readback(self._h, L.fptr(xv), L.fptr(rp), L.iptr(en),
                                               L.fptr(h0), L.fptr(h1)))
        return {"xv": xv, "rp": rp, "enabled": en, "h0": h0, "h1": h1}


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------
def unit_square_params(h: float = 0.02, gas_const: float = 35000.0, visc: float = 0.01,
                       wave_press: float = 20.0, couple: bool = True) -> L.Params2D:
    """The reference's 2D model in a unit square [0,1]^2 with support H = h:
    the box planes sit PARTICLE_RADIUS = H/4 inside the square (as the
    reference's left/right/ceiling planes), the wave-pressure lengths are
    the reference's scaled by 1/VIEW_HEIGHT (9.6), every particle stays
    enabled (threshold 0), and particles within h of the surface feed the
    wave (two-way coupling, kappa = 1)."""
    p = L.Params2D()
    prad = h / 4.0
    p.h = h
    p.rest_rho = 1000.0
    p.gas_const = gas_const
    p.visc = visc
    p.psi = psi_for(h, 1000.0)
    p.gravity[0], p.gravity[1] = 0.0, -9.8
    p.lo[0], p.lo[1] = 0.0, 0.0
    p.hi[0], p.hi[1] = 1.0, 1.0
    p.walls[0], p.walls[1], p.walls[2], p.walls[3] = prad, prad, 1.0 - prad, 1.0 - prad
    p.enable_vel = 0.0
    p.wave_ref = 0.5
    p.wave_press = wave_press
    p.wave_dy = 10.0 / 9.6
    p.wave_maskv = 0.2
    p.atten = 0.9995
    p.beta = 0.001
    p.couple_band = h if couple else 0.0
    p.couple_strength = 1.0 if couple else 0.0
    p.sph_substeps = 1
    p.check_finite = 0
    return p


def dam_break_block(n: int, spacing: float, origin, jitter: float, rng) -> np.ndarray:
    """n sites of a square lattice filling columns of a block from `origin`
    (x fastest, floor(0.5/s) sites per row, rows upward without bound), at
    site centres (i + 1/2) s, each jittered by U(-jitter s, +jitter s)."""
    ncol = max(1, int(np.floor(0.5 / spacing + 1e-4)))
    k = np.arange(n)
    i, j = k % ncol, k // ncol
    xy = np.stack([origin[0] + (i + 0.5) * spacing, origin[1] + (j + 0.5) * spacing], 1)
    xy += rng.uniform(-jitter * spacing, jitter * spacing, size=xy.shape)
    out = np.zeros((n, 4), np.float32)
    out[:, :2] = xy
    return out


def dam_break_batch(counts, lens, seed: int, dt: float = 4e-4, steps: int = 500,
                    wave_substeps: int = 10, c: float = 1.0, params: L.Params2D | None = None,
                    name: str = "") -> BatchSpec:
    """Scene s: counts[s] particles of a jittered (+-0.1 s) dam-break block of
    area 0.25 in the unit square (s = sqrt(0.25/n), mass rho0 s^2), and a
    wave column of lens[s] cells over x in [0, 1] started from Wave1D's
    InitWave, lambda = (c dt / (wave_substeps dx))^2."""
    p = unit_square_params() if params is None else params
    counts = np.asarray(counts, np.int64)
    lens = np.asarray(lens, np.int64)
    ns = counts.size
    prad = p.h / 4.0
    xv, h0s, h1s, mass, lam = [], [], [], [], []
    for s in range(ns):
        sp = float(np.sqrt(0.25 / counts[s]))
        rng = np.random.default_rng([seed, s])
        xv.append(dam_break_block(int(counts[s]), sp, (prad, prad), 0.1, rng))
        mass.append(np.float32(p.rest_rho) * np.float32(sp) * np.float32(sp))
        dx = (p.hi[0] - p.lo[0]) / float(lens[s])
        lam.append((c * dt / wave_substeps / dx) ** 2)
        a, b = wave_init_reference(int(lens[s]))
        h0s.append(a)
        h1s.append(b)
    return BatchSpec(params=p, n=counts, wave_len=lens,
                     wave_substeps=np.full(ns, wave_substeps, np.int64),
                     mass=np.array(mass, np.float32), lam=np.array(lam, np.float32),
                     xv=np.concatenate(xv), h0=np.concatenate(h0s), h1=np.concatenate(h1s),
                     dt=dt, steps=steps, name=name)


def batch_c3(nscenes: int = 256, seed: int = 2, steps: int = 500, dt: float = 4e-4) -> BatchSpec:
    """256 scenes, N_s = round(exp U[ln 1024, ln 65536]) and L_s =
    round(exp U[ln 256, ln 4096]) from one generator seeded `seed` (all N
    drawn first, then all L); h = 0.02; c = 1; 500 steps."""
    rng = np.random.default_rng(seed)
    counts = np.round(np.exp(rng.uniform(np.log(1024), np.log(65536), nscenes))).astype(np.int64)
    lens = np.round(np.exp(rng.uniform(np.log(256), np.log(4096), nscenes))).astype(np.int64)
    return dam_break_batch(counts, lens, seed, dt=dt, steps=steps,
                           name=f"batched 2D dam-break, {nscenes} scenes")


def small_batch(nscenes: int = 6, seed: int = 7, lo: int = 64, hi: int = 3000, steps: int = 20,
                dt: float = 4e-4, params: L.Params2D | None = None) -> BatchSpec:
    """A few uneven scenes for parity tests (sizes log-uniform in [lo, hi],
    columns in [16, 600])."""
    rng = np.random.default_rng(seed)
    counts = np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), nscenes))).astype(np.int64)
    lens = np.round(np.exp(rng.uniform(np.log(16), np.log(600), nscenes))).astype(np.int64)
    return dam_break_batch(counts, lens, seed, dt=dt, steps=steps, params=params,
                           name=f"small batch {nscenes}")


def reference_batch(nscenes: int = 1, steps: int = 100, shallow: bool = False) -> BatchSpec:
    """The SphWave2D application's scene (SphWave2D/Main.cpp:42,63-75 and
    InitParticle / init_grid, SphWaveKoschier2D_grid_cs.glsl:366-391): 4096
    particles on a 128 x 32 grid over (VIEW_WIDTH, 18 H), DT = 0.002, 2 SPH
    substeps; `nscenes` copies.  Wave: Wave1D with 1024 cells, 10 substeps,
    lambda = 0.01 (InitWaveEquation, Main.cpp:63-75), or with `shallow` the
    wave the application actually runs (InitShallowWaterEquation,
    Main.cpp:78-95,283): Shallow1D with 128 cells, one substep, lambda =
    0.001, dx = 0.1, started from its InitWave."""
    p = reference_params()
    prad, view, H = 0.025, 9.6, p.h
    n, cols = 4096, 128
    rows = n // cols
    k = np.arange(n)
    xy = np.stack([(k % cols) / cols, (k // cols) / rows], 1).astype(np.float32)
    xy = (np.array([view, 18.0 * H], np.float32) * xy).astype(np.float32)
    xy[:, 0] += np.float32(prad)
    xy[:, 1] += np.float32(0.5 * view - 15.0 * prad + prad)
    xv = np.zeros((n, 4), np.float32)
    xv[:, :2] = xy
    if shallow:
        Lw, nsub, lam, model = 128, 1, 0.001, L.CWA2D_WAVE_SHALLOW
        h0, h1 = wave_init_shallow(Lw, view)
    else:
        Lw, nsub, lam, model = 1024, 10, 0.01, L.CWA2D_WAVE_LEAPFROG
        h0, h1 = wave_init_reference(Lw)
    mass = np.float32(2 * prad) * np.float32(2 * prad) * np.float32(1000.0)
    return BatchSpec(params=p, n=np.full(nscenes, n, np.int64),
                     wave_len=np.full(nscenes, Lw, np.int64),
                     wave_substeps=np.full(nscenes, nsub, np.int64),
                     mass=np.full(nscenes, mass, np.float32),
                     lam=np.full(nscenes, lam, np.float32), xv=np.tile(xv, (nscenes, 1)),
                     h0=np.tile(h0, nscenes), h1=np.tile(h1, nscenes), dt=0.002, steps=steps,
                     name="SphWave2D reference scene" + (" (Shallow1D)" if shallow else ""),
                     wave_model=np.full(nscenes, model, np.int64),
                     wave_dx=np.full(nscenes, 0.1, np.float32))
