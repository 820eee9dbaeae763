"""Scene builders: the reference's own scene and the benchmark workloads.

Each builder returns a ``SceneSpec`` (parameters, initial particle arrays,
heightfield, dt, steps) that both the HIP library and the CPU oracle
consume, so both start from identical host arrays.

Choices that the workload descriptions leave open are made here and listed
in each docstring.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .scene import reference_params


@dataclass
class SceneSpec:
    name: str
    params: L.Params
    pos: np.ndarray               # (n, 4) float32
    vel: np.ndarray               # (n, 4) float32
    wave_size: tuple              # (W, H)
    height0: np.ndarray           # (H, W) float32
    height1: np.ndarray           # (H, W) float32
    dt: float
    steps: int
    info: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return self.pos.shape[0]

    def make_scene(self):
        from .scene import Scene
        return Scene(self.params, self.pos, self.vel, None, self.wave_size, self.height0,
                     self.height1)


def lattice(n, lo, hi, spacing, site_offset, jitter, order, seed) -> np.ndarray:
    """cwa_gen_lattice: jittered lattice, axis order[0] fastest."""
    out = np.zeros((n, 4), np.float32)
    lo_a = (C.c_float * 3)(*lo)
    hi_a = (C.c_float * 3)(*hi)
    ord_a = (C.c_int32 * 3)(*order)
    rc = L.lib().cwa_gen_lattice(L.fptr(out), n, lo_a, hi_a, spacing, site_offset, jitter,
                                 ord_a, seed)
    if rc != n:
        raise L.CwaError(rc, "lattice generation failed")
    return out


def bumps(w, h, nbumps, amp, sigma, seed) -> np.ndarray:
    """cwa_gen_bumps: sum of Gaussian bumps on a w x h texel grid."""
    out = np.zeros((h, w), np.float32)
    L.check(L.lib().cwa_gen_bumps(L.fptr(out), w, h, nbumps, amp[0], amp[1], sigma[0],
                                  sigma[1], seed))
    return out


def grid_for_box(p: L.Params, lo, hi, cell, ydiv: int = 1) -> None:
    """Uniform grid over [lo, hi] with the given cell size (cell >= h) along
    x and z and cell / ydiv along y."""
    p.cell_size = cell
    p.grid_ydiv = ydiv
    cy = float(np.float32(cell) / np.float32(ydiv))
    for a in range(3):
        p.grid_min[a] = lo[a]
        edge = cy if a == 1 else cell
        p.grid_dims[a] = max(1, math.ceil((hi[a] - lo[a]) / edge - 1e-4))


def _textbook_params(h, spacing, box_lo, box_hi, wave_w, wave_h, dt, c, atten) -> L.Params:
    """The physical constants shared by the dam-break workloads.

    Mueller poly6 / spiky (45/(pi h^6)) / viscosity laplacian (45/(pi h^6)),
    stiffness k = 3, rest density 1000, mu = 0.1, g = -9.8; mass = rho0 * s^3;
    particle radius s/2.  The reference's torque term (force_comp.glsl:103-104)
    is switched off: it feeds the previous force back scaled by |pos|, which
    grows without bound once |pos| > 1 (outside the reference's half-unit
    box).  Crest, foam, drag and wave terms keep the reference's values.
    The heightfield spans the box's xz face; lambda = (c dt / dx)^2 per axis.
    Two-way coupling inside the band |y - H| < h with unit strength.
    """
    p = reference_params()
    p.mass = 1000.0 * spacing ** 3
    p.h = h
    p.visc = 0.1
    p.rest_rho = 1000.0
    p.gas_const = 3.0
    p.particle_radius = 0.5 * spacing
    p.spiky_coeff = 45.0
    p.visc_coeff = 45.0
    p.gravity[:] = [0.0, -9.8, 0.0]
    p.torque_coeff = 0.0
    for a in range(3):
        p.lower[a] = box_lo[a]
        p.upper[a] = box_hi[a]
    p.wave_origin[:] = [box_lo[0], box_lo[2]]
    p.wave_uv_scale[:] = [1.0 / (box_hi[0] - box_lo[0]), 1.0 / (box_hi[2] - box_lo[2])]
    dx = (box_hi[0] - box_lo[0]) / wave_w
    dz = (box_hi[2] - box_lo[2]) / wave_h
    p.lambda_x = (c * dt / dx) ** 2
    p.lambda_z = (c * dt / dz) ** 2
    p.atten = atten
    p.beta = 0.0
    p.wave_type = 1.0
    p.couple_band = h
    p.couple_strength = 1.0
    grid_for_box(p, box_lo, box_hi, h)
    return p


def reference_scene(steps: int = 200) -> SceneSpec:
    """The reference app's own scene: 20 480 particles from make_cube
    (Main.cpp:735-759: 64 x 5 x 64 lattice, spacing 2*0.85*0.005, sites at
    i*spacing, z fastest then y then x), zero velocity, 64 x 64 heightfield
    with the splash initialisation (wave_comp.glsl:88-95) in both history
    levels, dt = 5e-5 (the shaders' constant)."""
    p = reference_params()
    s = 2.0 * 0.85 * 0.005
    pos = lattice(20480, (0.0, 0.0, 0.0), (64 * s, 5 * s, 64 * s), s, 0.0, 0.0, (2, 1, 0), 0)
    W = H = 64
    # splash InitWave, computed here in closed form (the scene's own
    # wave_reinit() produces the same field on the device)
    ii, jj = np.meshgrid(np.arange(W), np.arange(H))
    c0 = (int(0.25 * W), int(0.25 * H))
    c1 = (int(0.75 * W), int(0.75 * H))
    d = np.minimum(np.hypot(ii - c0[0], jj - c0[1]), np.hypot(ii - c1[0], jj - c1[1]))
    t = np.clip((d - 5.0) / (0.0 - 5.0), 0.0, 1.0)
    h0 = (0.5 * t * t * (3.0 - 2.0 * t)).astype(np.float32)
    return SceneSpec("reference", p, pos, np.zeros_like(pos), (W, H), h0, h0.copy(), 5e-5,
                     steps, {"init": "make_cube + splash"})


def dam_break_small(seed: int = 0, steps: int = 200) -> SceneSpec:
    """Plain coupled 3D step, 4 096 particles + 128^2 heightfield.

    Unit box; block [0,0.3]x[0,0.6]x[0,0.3]; lattice spacing 0.02, sites at
    lo + (i + 1/2) s, jitter +-0.1 s, enumerated z fastest, then x, then y
    from the floor, the first 4 096 kept.  h = 0.04, dt = 1e-3, c = 1,
    damping (attenuation) 0.999, zero initial height."""
    s, h, dt = 0.02, 0.04, 1e-3
    n, W, H = 4096, 128, 128
    pos = lattice(n, (0.0, 0.0, 0.0), (0.3, 0.6, 0.3), s, 0.5, 0.1, (2, 0, 1), seed)
    p = _textbook_params(h, s, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), W, H, dt, 1.0, 0.999)
    z = np.zeros((H, W), np.float32)
    return SceneSpec("dam_break_4k_128", p, pos, np.zeros_like(pos), (W, H), z, z.copy(), dt,
                     steps, {"spacing": s})


def dam_break_large(seed: int = 1, steps: int = 1000, n: int = 1 << 20,
                    wave: int = 2048, stiffness: float = 1000.0, dt: float = 1e-4,
                    ydiv: int = 2) -> SceneSpec:
    """Plain coupled 3D step at scale: 1 048 576 particles + 2048^2 heightfield.

    Box [0,2]x[0,1]x[0,1]; block = the x-prefix slab holding 40% of the box
    volume; spacing s = (0.8/n)^(1/3); sites z fastest, then y, then x, the
    slowest axis extended until n sites exist (x reaches ~0.81); jitter
    +-0.1 s; h = 1.95 s so that (4/3) pi (h/s)^3 ~ 31 neighbours; grid cell = h.
    Heightfield: 8 Gaussian bumps, amplitude U[0.01, 0.05], sigma U[5, 40]
    texels, bump seed = seed + 0x5EED; both history levels equal.
    Stiffness k = 1000 (sound speed ~32 m/s, ~10x the dam's fall speed:
    the weakly compressible regime, so the column stays near rest density
    and keeps ~30 neighbours per particle; the small workload's k = 3 would
    compress a metre-high column about 4x).  Grid: cell = h along x and z,
    h / ydiv along y (ydiv = 2, the fastest of 1..8 on MI355X), so the
    neighbour search keeps only the rows each column's part of the support
    sphere reaches.  c = 1 and dt = 1e-4 (dt = 1e-3
    would break the wave CFL limit on a 2048-texel unit axis; 1e-4 keeps
    the SPH acoustic CFL number at ~0.18); attenuation 0.999."""
    s = (0.8 / n) ** (1.0 / 3.0)
    h = 1.95 * s
    W = H = wave
    pos = lattice(n, (0.0, 0.0, 0.0), (0.8, 1.0, 1.0), s, 0.5, 0.1, (2, 1, 0), seed)
    p = _textbook_params(h, s, (0.0, 0.0, 0.0), (2.0, 1.0, 1.0), W, H, dt, 1.0, 0.999)
    grid_for_box(p, (0.0, 0.0, 0.0), (2.0, 1.0, 1.0), h, ydiv)
    p.gas_const = stiffness
    h0 = bumps(W, H, 8, (0.01, 0.05), (5.0, 40.0), seed + 0x5EED)
    return SceneSpec(f"dam_break_{n}_{wave}", p, pos, np.zeros_like(pos), (W, H), h0,
                     h0.copy(), dt, steps, {"spacing": s})


def dam_break_diff(n: int = 65536, wave: int = 512, seed: int = 3, dt: float = 5e-4,
                   steps: int = 256) -> SceneSpec:
    """Differentiable coupled 3D step: 65 536 particles + 512^2 heightfield.

    The small workload's physics (unit box, block [0,0.3]x[0,0.6]x[0,0.3],
    k = 3, rho0 = 1000, mu = 0.1, h = 2 s) at n particles with spacing
    s = (0.054/n)^(1/3), sites z fastest, then x, then y, jitter +-0.1 s.
    Heightfield: 4 Gaussian bumps, amplitude U[0.01, 0.05], sigma
    U[wave/64, wave/16] texels (bump seed = seed + 0x5EED), both history
    levels equal; c = 1, attenuation 0.999; dt = 5e-4 keeps the 512-texel
    wave inside its CFL limit.  The gradient check runs the same builder at
    n = 4096, 64^2 texels, 32 steps."""
    s = (0.3 * 0.6 * 0.3 / n) ** (1.0 / 3.0)
    h = 2.0 * s
    W = H = wave
    pos = lattice(n, (0.0, 0.0, 0.0), (0.3, 0.6, 0.3), s, 0.5, 0.1, (2, 0, 1), seed)
    p = _textbook_params(h, s, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), W, H, dt, 1.0, 0.999)
    h0 = bumps(W, H, 4, (0.01, 0.05), (wave / 64.0, wave / 16.0), seed + 0x5EED)
    return SceneSpec(f"dam_break_diff_{n}_{wave}", p, pos, np.zeros_like(pos), (W, H), h0,
                     h0.copy(), dt, steps, {"spacing": s, "wave_speed": 1.0})
