"""Batched parameter fits of the coupled 3D step (cwa_fit_*): many scenes,
each fitting its six physical scalars to a target heightfield by gradient
descent, every gradient evaluation one replay of one hipGraph that holds
all scenes' checkpointed forward passes, losses and adjoint sweeps.

    specs, targets, truth = fit.fit_workload(nscenes=64)   # seeded targets
    with fit.FitRun(specs, targets) as run:
        result = run.descend(iters=50)
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from . import configs
from .gradient import Rollout
from .scene import Scene

THETA_NAMES = ("visc", "gas_const", "rest_rho", "wave_speed", "atten", "couple_strength")


class FitBatch:
    """cwa_fit over existing scenes: run() -> (losses, list of gradient dicts)."""

    def __init__(self, scenes: list[Scene], targets: list[np.ndarray], steps: int,
                 ckpt_every: int, dt: float):
        self._lib = L.lib()
        self.ns = len(scenes)
        self._targets = [np.ascontiguousarray(t, np.float32) for t in targets]
        hs = (C.c_void_p * self.ns)(*[s._h.value for s in scenes])
        tp = (C.POINTER(C.c_float) * self.ns)(*[L.fptr(t) for t in self._targets])
        self._h = C.c_void_p()
        L.check(self._lib.cwa_fit_create(self.ns, hs, tp, int(steps), int(ckpt_every),
                                         float(dt), C.byref(self._h)))

    def run(self):
        losses = (C.c_double * self.ns)()
        grads = (L.Grad * self.ns)()
        L.check(self._lib.cwa_fit_run(self._h, losses, grads))
        return (np.array(losses[:]),
                [{k: getattr(g, k) for k, _ in L.Grad._fields_} for g in grads])

    def tape_bytes(self, s: int) -> int:
        b = C.c_int64()
        L.check(self._lib.cwa_fit_tape_bytes(self._h, int(s), C.byref(b)))
        return int(b.value)

    def close(self):
        if self._h:
            L.check(self._lib.cwa_fit_destroy(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fit_workload(nscenes: int = 64, seed: int = 4, steps: int = 128, dt: float = 3e-4,
                 n_range=(8192, 131072), side_range=(128, 1024)):
    """Scene s: N_s = round(exp U[ln 8k, ln 128k]) particles and a square
    heightfield of side U{128..1024} texels (one generator seeded `seed`:
    all N, then all sides, then the 6 x nscenes target factors), the
    differentiable dam-break physics (configs.dam_break_diff) at dt = 3e-4
    (inside the wave CFL limit at 1024 texels), `steps` steps.  Targets:
    the same scene run on the device with its six scalars scaled by
    U[0.8, 1.2] (wave speed through lambda, attenuation capped at 0.9999).
    Returns (specs, targets, true factors)."""
    rng = np.random.default_rng(seed)
    ns = np.round(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]), nscenes)))
    sides = rng.integers(side_range[0], side_range[1] + 1, nscenes)
    factors = rng.uniform(0.8, 1.2, (nscenes, 6))
    specs, targets = [], []
    for s in range(nscenes):
        spec = configs.dam_break_diff(n=int(ns[s]), wave=int(sides[s]), seed=seed * 1000 + s,
                                      dt=dt, steps=steps)
        specs.append(spec)
        ro = Rollout(spec, steps, 16)
        th = np.array(ro.theta0(1.0)) * factors[s]
        th[4] = min(th[4], 0.9999)
        p = ro.params_for(th)
        with Scene(p, spec.pos, spec.vel, None, spec.wave_size, spec.height0,
                   spec.height1) as sc:
            sc.step(spec.dt, steps)
            targets.append(sc.readback()["height0"].copy())
    return specs, targets, factors


@dataclass
class FitResult:
    losses: np.ndarray          # (iters + 1, nscenes): loss before each update, and final
    u: np.ndarray               # (iters + 1, nscenes, 6): theta / theta0
    eval_s: list                # wall time of each graph evaluation
    tape_bytes: np.ndarray      # per scene
    scene_bytes: np.ndarray     # per scene
    best_loss: np.ndarray       # per scene: lowest loss evaluated
    best_u: np.ndarray          # per scene: the u that reached it


class FitRun:
    """Scenes at their initial guesses (theta0 of each spec), the fit graph,
    and an Adam descent in u = theta / theta0."""

    def __init__(self, specs, targets, ckpt_every: int = 16):
        import torch  # noqa: F401  (device memory accounting below)

        self.specs = specs
        self.ro = [Rollout(sp, sp.steps, ckpt_every) for sp in specs]
        self.theta0 = np.array([r.theta0(1.0) for r in self.ro])
        self.scenes = []
        self.scene_bytes = []
        for sp, r, th in zip(specs, self.ro, self.theta0):
            f0 = _free_bytes()
            self.scenes.append(Scene(r.params_for(th), sp.pos, sp.vel, None, sp.wave_size,
                                     sp.height0, sp.height1))
            self.scene_bytes.append(max(0, f0 - _free_bytes()))
        dts = {sp.dt for sp in specs}
        steps = {sp.steps for sp in specs}
        if len(dts) != 1 or len(steps) != 1:
            raise ValueError("one dt and one step count for the whole batch")
        self.fit = FitBatch(self.scenes, targets, steps.pop(), ckpt_every, dts.pop())

    def set_u(self, u: np.ndarray) -> None:
        for s, (r, th0) in enumerate(zip(self.ro, self.theta0)):
            th = th0 * u[s]
            th[4] = min(th[4], 0.99999)
            self.scenes[s].set_params(r.params_for(th))

    def evaluate(self, u: np.ndarray):
        """Losses and d loss / d u (nscenes x 6) at u."""
        self.set_u(u)
        losses, grads = self.fit.run()
        gu = np.array([r.chain_theta(g, th0 * uu) * th0
                       for r, g, th0, uu in zip(self.ro, grads, self.theta0, u)])
        return losses, gu

    def descend(self, iters: int = 50, lr: float = 0.01, beta1: float = 0.9,
                beta2: float = 0.999) -> FitResult:
        ns = len(self.scenes)
        u = np.ones((ns, 6))
        m = np.zeros_like(u)
        v = np.zeros_like(u)
        L_hist, u_hist, times = [], [u.copy()], []
        for it in range(iters + 1):
            t0 = time.perf_counter()
            loss, g = self.evaluate(u)
            times.append(time.perf_counter() - t0)
            L_hist.append(loss)
            if it == iters:
                break
            m = beta1 * m + (1 - beta1) * g
            v = beta2 * v + (1 - beta2) * g * g
            mh = m / (1 - beta1 ** (it + 1))
            vh = v / (1 - beta2 ** (it + 1))
            u = np.clip(u - lr * mh / (np.sqrt(vh) + 1e-12), 0.5, 1.5)
            u_hist.append(u.copy())
        Lh, uh = np.array(L_hist), np.array(u_hist)
        best = np.argmin(Lh, axis=0)
        return FitResult(Lh, uh, times, np.array([self.fit.tape_bytes(s) for s in range(ns)]),
                         np.array(self.scene_bytes), Lh[best, np.arange(ns)],
                         uh[best, np.arange(ns)])

    def close(self):
        self.fit.close()
        for s in self.scenes:
            s.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _free_bytes() -> int:
    import torch

    return int(torch.cuda.mem_get_info()[0])
