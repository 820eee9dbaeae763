"""Gradients of the coupled step: the tape binding and a torch.autograd
Function over it.

The physics runs in the HIP library (checkpointed forward, hand-written
adjoint kernels); torch only carries the host-side tensors through
autograd, so a loss built in torch from the final heightfield (and/or the
final particle positions) can be differentiated with respect to the six
physical scalars and the initial particle velocities.

    theta = torch.tensor([visc, gas_const, rest_rho, wave_speed, atten,
                          couple_strength], dtype=torch.float64,
                         requires_grad=True)
    height, pos = CoupledRollout.apply(theta, vel0, rollout)
    loss = ((height - target) ** 2).mean()
    loss.backward()            # theta.grad, vel0.grad
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .scene import Scene

PARAM_NAMES = ("visc", "gas_const", "rest_rho", "wave_speed", "atten", "couple_strength")


class Tape:
    """cwa_tape: `nsteps` steps of `scene` from its current state with a
    checkpoint every `ckpt_every` steps."""

    def __init__(self, scene: Scene, nsteps: int, ckpt_every: int):
        self.scene = scene
        self._lib = L.lib()
        self._t = C.c_void_p()
        L.check(self._lib.cwa_tape_create(scene._h, int(nsteps), int(ckpt_every),
                                          C.byref(self._t)))
        self.n = scene.particles().n
        _, self.W, self.H, _ = scene.height(0)

    def forward(self, dt: float) -> None:
        L.check(self._lib.cwa_tape_forward(self._t, float(dt)))

    def backward(self, dheight0=None, dheight1=None, dpos=None, dvel=None):
        """Seeds are derivatives of the loss with respect to the final state
        (heights (H, W); particle arrays (n, 3) or (n, 4), original order).
        Returns (cwa_grad as dict, dpos0 (n, 3), dvel0 (n, 3))."""
        def field(a):
            return None if a is None else np.ascontiguousarray(a, np.float32).reshape(self.H,
                                                                                        self.W)

        def parts(a):
            if a is None:
                return None
            a = np.asarray(a, np.float32).reshape(self.n, -1)
            out = np.zeros((self.n, 4), np.float32)
            out[:, :3] = a[:, :3]
            return out

        seeds = [field(dheight0), field(dheight1), parts(dpos), parts(dvel)]
        g = L.Grad()
        dp0 = np.zeros((self.n, 4), np.float32)
        dv0 = np.zeros((self.n, 4), np.float32)
        L.check(self._lib.cwa_tape_backward(self._t, *[L.fptr(s) for s in seeds], C.byref(g),
                                            L.fptr(dp0), L.fptr(dv0)))
        grad = {k: getattr(g, k) for k, _ in L.Grad._fields_}
        return grad, dp0[:, :3], dv0[:, :3]

    def close(self) -> None:
        if self._t:
            L.check(self._lib.cwa_tape_destroy(self._t))
            self._t = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Rollout:
    """A differentiable run: a workload (SceneSpec), the number of steps and
    the checkpoint interval.  The wave speed enters through
    lambda = (c dt / dx)^2 per axis."""

    spec: object
    steps: int
    ckpt_every: int = 16

    def texel(self):
        p = self.spec.params
        W, H = self.spec.wave_size
        return 1.0 / (W * p.wave_uv_scale[0]), 1.0 / (H * p.wave_uv_scale[1])

    def params_for(self, theta) -> L.Params:
        visc, k, rho0, c, atten, kappa = [float(x) for x in theta]
        p = self.spec.params.copy()
        dx, dz = self.texel()
        p.visc, p.gas_const, p.rest_rho = visc, k, rho0
        p.lambda_x = (c * self.spec.dt / dx) ** 2
        p.lambda_z = (c * self.spec.dt / dz) ** 2
        p.atten, p.couple_strength = atten, kappa
        return p

    def theta0(self, wave_speed: float):
        p = self.spec.params
        return [p.visc, p.gas_const, p.rest_rho, wave_speed, p.atten, p.couple_strength]

    def run(self, theta, vel0):
        """Forward with checkpoints; returns (tape, height, pos) with the
        final heightfield (H, W) and positions (n, 3)."""
        p = self.params_for(theta)
        vel = np.zeros((self.spec.n, 4), np.float32)
        vel[:, :3] = np.asarray(vel0, np.float32).reshape(-1, 3)
        scene = Scene(p, self.spec.pos, vel, None, self.spec.wave_size, self.spec.height0,
                      self.spec.height1)
        tape = Tape(scene, self.steps, self.ckpt_every)
        tape.forward(self.spec.dt)
        st = scene.readback()
        return tape, st["height0"], st["pos"][:, :3]

    def chain_theta(self, grad: dict, theta) -> np.ndarray:
        """d loss / d theta from the library's parameter gradients."""
        c = float(theta[3])
        dx, dz = self.texel()
        dt = self.spec.dt
        lx, lz = (c * dt / dx) ** 2, (c * dt / dz) ** 2
        dc = (grad["lambda_x"] * 2.0 * lx + grad["lambda_z"] * 2.0 * lz) / c if c != 0 else 0.0
        return np.array([grad["visc"], grad["gas_const"], grad["rest_rho"], dc, grad["atten"],
                         grad["couple_strength"]], np.float64)


def _autograd_function():
    import torch

    class CoupledRollout(torch.autograd.Function):
        """(theta (6,), vel0 (n, 3)) -> (final heightfield (H, W), final
        positions (n, 3)); backward runs the library's adjoint."""

        @staticmethod
        def forward(ctx, theta, vel0, rollout: Rollout):
            th = theta.detach().cpu().numpy().astype(np.float64)
            tape, height, pos = rollout.run(th, vel0.detach().cpu().numpy())
            ctx.tape, ctx.rollout, ctx.theta = tape, rollout, th
            ctx.dtypes = (theta.dtype, vel0.dtype)
            return torch.from_numpy(height.copy()), torch.from_numpy(pos.copy())

        @staticmethod
        def backward(ctx, g_height, g_pos):
            gh = None if g_height is None else g_height.detach().cpu().numpy()
            gp = None if g_pos is None else g_pos.detach().cpu().numpy()
            grad, _, dv0 = ctx.tape.backward(dheight0=gh, dpos=gp)
            dtheta = ctx.rollout.chain_theta(grad, ctx.theta)
            scene = ctx.tape.scene
            ctx.tape.close()
            scene.close()
            return (torch.from_numpy(dtheta).to(ctx.dtypes[0]),
                    torch.from_numpy(dv0.copy()).to(ctx.dtypes[1]), None)

    return CoupledRollout


_FN = None


def coupled_rollout(theta, vel0, rollout: Rollout):
    """torch entry point: CoupledRollout.apply(theta, vel0, rollout)."""
    global _FN
    if _FN is None:
        _FN = _autograd_function()
    return _FN.apply(theta, vel0, rollout)
