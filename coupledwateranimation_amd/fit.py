"""Batched parameter fits of the coupled 3D step (cwa_fit_*): many scenes,
each fitting its six physical scalars to a target heightfield by gradient
descent, every gradient evaluation one replay of one hipGraph that holds
all scenes' checkpointed forward passes, losses and adjoint sweeps.

    specs, targets, truth = fit.fit_workload(nscenes=64)   # seeded targets
    with fit.FitRun(specs, targets) as run:
        result = run.descend(iters=50)      # L-BFGS per scene
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
    U[0.8, 1.2] (wave speed through lambda; for the attenuation the factor
    scales the per-step loss 1 - atten, so the target's wave keeps about the
    initial guess's 0.999^128 = 0.88 of its amplitude and stays in view).
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
        th = _scaled(np.array(ro.theta0(1.0)), factors[s])
        p = ro.params_for(th)
        with Scene(p, spec.pos, spec.vel, None, spec.wave_size, spec.height0,
                   spec.height1) as sc:
            sc.step(spec.dt, steps)
            targets.append(sc.readback()["height0"].copy())
    return specs, targets, factors


def _scaled(theta: np.ndarray, factors: np.ndarray) -> np.ndarray:
    th = theta * factors
    th[..., ATTEN] = 1.0 - (1.0 - theta[..., ATTEN]) * factors[..., ATTEN]
    return th


# Fit coordinates: theta = theta0 * x for the multiplicative scalars (visc,
# gas_const, rest_rho, wave_speed, couple_strength); the attenuation as
# x = log(1 - atten), so the steep approach to atten -> 1 is a straight line
# and the bound atten < 1 holds without a clamp inside the descent.
ATTEN = 4
X_LO = np.array([0.5, 0.5, 0.5, 0.5, np.log(1e-5), 0.5])
X_HI = np.array([1.5, 1.5, 1.5, 1.5, np.log(0.5), 1.5])
# Scale of each fit coordinate in the descent: the spread of the scalars
# around their guesses (a fifth; for log(1 - atten) a fifth too).
X_SCALE = np.full(6, 0.2)


@dataclass
class FitResult:
    losses: np.ndarray          # (iters + 1, nscenes): loss of each evaluated iterate
    accepted: np.ndarray        # (iters + 1, nscenes): loss of the accepted iterate
    x: np.ndarray               # (iters + 1, nscenes, 6): accepted fit coordinates
    eval_s: list                # wall time of each graph evaluation
    tape_bytes: np.ndarray      # per scene
    scene_bytes: np.ndarray     # per scene
    best_loss: np.ndarray       # per scene: the final accepted loss (the lowest evaluated)
    best_theta: np.ndarray      # per scene: the physical scalars that reached it
    rejected: np.ndarray        # per scene: proposals rejected


class FitRun:
    """Scenes at their initial guesses (theta0 of each spec), the fit graph,
    and a quasi-Newton descent with per-scene step control in the fit
    coordinates (see X_LO / X_HI / X_SCALE)."""

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

    def x0(self) -> np.ndarray:
        """Fit coordinates of the initial guesses."""
        x = np.ones((len(self.scenes), 6))
        x[:, ATTEN] = np.log1p(-self.theta0[:, ATTEN])
        return x

    def theta_of(self, x: np.ndarray) -> np.ndarray:
        th = self.theta0 * x
        th[:, ATTEN] = -np.expm1(x[:, ATTEN])
        return th

    def evaluate(self, x: np.ndarray):
        """Losses and d loss / d x (nscenes x 6) at fit coordinates x."""
        th = self.theta_of(x)
        for s, r in enumerate(self.ro):
            self.scenes[s].set_params(r.params_for(th[s]))
        losses, grads = self.fit.run()
        gth = np.array([r.chain_theta(g, t) for r, g, t in zip(self.ro, grads, th)])
        gx = gth * self.theta0
        gx[:, ATTEN] = gth[:, ATTEN] * -(1.0 - th[:, ATTEN])   # d atten / d log(1 - atten)
        return losses, gx

    def descend(self, iters: int = 50, memory: int = 5, first_step: float = 0.25,
                scale=None) -> FitResult:
        """Limited-memory BFGS per scene in scaled fit coordinates z = (x -
        x0) / scale, one graph evaluation per iteration.  The first direction
        is the gradient, scaled so its largest component moves `first_step`;
        later ones come from the scene's last `memory` accepted (s, y) pairs
        (pairs with s.y <= 0 are skipped).  A proposal whose loss is not
        below the accepted one is rejected and the step along the same
        direction halved; an accepted one resets the step to 1 and takes a
        new direction, so the accepted loss never rises."""
        ns = len(self.scenes)
        sc = np.asarray(X_SCALE if scale is None else scale, np.float64)
        x0 = self.x0()
        z = np.zeros_like(x0)
        t0 = time.perf_counter()
        f, gx = self.evaluate(x0)
        times = [time.perf_counter() - t0]
        g = gx * sc
        hist = [[] for _ in range(ns)]
        alpha = np.ones(ns)
        rejected = np.zeros(ns, np.int64)

        def direction(k):
            q = -g[k].copy()
            pairs = hist[k]
            a = []
            for s_, y_ in reversed(pairs):
                rho = 1.0 / float(y_ @ s_)
                ai = rho * float(s_ @ q)
                q -= ai * y_
                a.append((rho, ai))
            if pairs:
                s_, y_ = pairs[-1]
                q *= float(s_ @ y_) / float(y_ @ y_)
            else:
                q *= first_step / max(np.abs(q).max(), 1e-300)
            for (s_, y_), (rho, ai) in zip(pairs, reversed(a)):
                b = rho * float(y_ @ q)
                q += (ai - b) * s_
            return q

        d = np.array([direction(k) for k in range(ns)])
        lo, hi = (X_LO - x0) / sc, (X_HI - x0) / sc
        L_hist, A_hist, x_hist = [f.copy()], [f.copy()], [x0.copy()]
        for _ in range(iters):
            zp = np.clip(z + alpha[:, None] * d, lo, hi)
            t1 = time.perf_counter()
            fp, gpx = self.evaluate(x0 + zp * sc)
            times.append(time.perf_counter() - t1)
            gp = gpx * sc
            ok = np.isfinite(fp) & (fp < f) & np.all(np.isfinite(gp), axis=1)
            for k in np.nonzero(ok)[0]:
                s_, y_ = zp[k] - z[k], gp[k] - g[k]
                if float(s_ @ y_) > 1e-12 * np.linalg.norm(s_) * np.linalg.norm(y_):
                    hist[k].append((s_, y_))
                    del hist[k][:-memory]
                z[k], f[k], g[k] = zp[k], fp[k], gp[k]
                alpha[k] = 1.0
                d[k] = direction(k)
            alpha[~ok] *= 0.5
            rejected += ~ok
            L_hist.append(fp)
            A_hist.append(f.copy())
            x_hist.append(x0 + z * sc)
        x = x0 + z * sc
        return FitResult(np.array(L_hist), np.array(A_hist), np.array(x_hist), times,
                         np.array([self.fit.tape_bytes(s) for s in range(ns)]),
                         np.array(self.scene_bytes), f.copy(), self.theta_of(x), rejected)

    def close(self):
        self.fit.close()
        for s in self.scenes:
            s.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def true_theta(run: FitRun, factors: np.ndarray) -> np.ndarray:
    """The physical scalars the targets were made with (fit_workload)."""
    return _scaled(run.theta0, factors)


def theta_errors(theta: np.ndarray, truth: np.ndarray) -> np.ndarray:
    """Per scene and scalar: |theta / truth - 1| for the multiplicative
    scalars, |(1 - atten) / (1 - atten_true) - 1| for the attenuation."""
    e = np.abs(theta / truth - 1.0)
    e[:, ATTEN] = np.abs((1.0 - theta[:, ATTEN]) / (1.0 - truth[:, ATTEN]) - 1.0)
    return e


def _free_bytes() -> int:
    import torch

    return int(torch.cuda.mem_get_info()[0])
