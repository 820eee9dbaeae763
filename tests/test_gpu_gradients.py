"""Gradient path: the HIP adjoint (fp32) against FP64 central finite
differences of the oracle.

Parameters are stored as fp32 in cwa_params, so each finite difference uses
the perturbed values as actually stored, theta+- = fl32(theta (1 +- eps)),
and divides by their exact difference.  Particle perturbations are applied
to the FP64 oracle's own (double) arrays.

Bars:
  one-step / few-step seeds on every output ... 2e-4 relative (per scalar)
  4 096-particle, 64^2, 32-step MSE loss ....... 1e-3 relative, six scalars
                                                 and velocity directions
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS = 1e-6
SCALARS = ("visc", "gas_const", "rest_rho", "lambda_x", "lambda_z", "atten", "couple_strength")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _oracle_final(oracle, spec, params, steps, pos=None, vel=None):
    pos = spec.pos.astype(np.float64) if pos is None else pos
    vel = spec.vel.astype(np.float64) if vel is None else vel
    o = oracle.OracleScene(params, pos, vel, None, spec.wave_size, spec.height0, spec.height1,
                           double=True, threads=8)
    o.step(spec.dt, steps)
    st = o.state()
    o.close()
    return st


def _fd_scalar(oracle, spec, steps, name, loss_of_state):
    p0 = spec.params
    v = float(getattr(p0, name))
    pp, pm = p0.copy(), p0.copy()
    setattr(pp, name, v * (1 + EPS))
    setattr(pm, name, v * (1 - EPS))
    step = float(getattr(pp, name)) - float(getattr(pm, name))   # as stored (fp32)
    assert step != 0.0
    lp = loss_of_state(_oracle_final(oracle, spec, pp, steps))
    lm = loss_of_state(_oracle_final(oracle, spec, pm, steps))
    return (lp - lm) / step


def _fd_direction(oracle, spec, steps, which, d, loss_of_state, scale):
    base_p = spec.pos.astype(np.float64)
    base_v = spec.vel.astype(np.float64)
    h = EPS * scale
    out = []
    for sgn in (1.0, -1.0):
        p, v = base_p.copy(), base_v.copy()
        tgt = p if which == "pos" else v
        tgt[:, :3] += sgn * h * d
        out.append(loss_of_state(_oracle_final(oracle, spec, spec.params, steps, p, v)))
    return (out[0] - out[1]) / (2 * h)


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


@pytest.mark.parametrize("couple,ydiv,n,wave", [(True, 1, 256, 16), (False, 1, 256, 16),
                                                (True, 4, 256, 16), (True, 1, 4096, 256)])
def test_few_step_adjoint_all_seeds(gpu, cwa, oracle, couple, ydiv, n, wave):
    """Loss = <a, h^n> + <b, h^(n-1)> + <c, x> + <d, v> after 3 steps of a
    small scene: parameter gradients and directional derivatives in
    the initial positions and velocities against FP64 FD.  The 4 096-particle
    case on 256^2 texels sends the heightfield taps of the local adjoint
    through its LDS windows (the particles cover texels 0..77) and its
    last-column / last-row strips (the wave normal's taps one texture width
    away)."""
    from coupledwateranimation_amd import configs
    from coupledwateranimation_amd.gradient import Tape

    spec = configs.dam_break_diff(n=n, wave=wave, dt=2e-3, steps=3)
    if ydiv != 1:
        configs.grid_for_box(spec.params, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), spec.params.h, ydiv)
    if not couple:
        spec.params.couple_band = 0.0
    rng = np.random.default_rng(5)
    W, H = spec.wave_size
    a, b = rng.normal(size=(H, W)), rng.normal(size=(H, W))
    c, d = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    rng2 = np.random.default_rng(6)
    spec.vel[:, :3] = rng2.normal(scale=0.2, size=(n, 3)).astype(np.float32)

    def loss(st):
        return float((a * st["height0"]).sum() + (b * st["height1"]).sum() +
                     (c * st["pos"][:, :3]).sum() + (d * st["vel"][:, :3]).sum())

    scene = spec.make_scene()
    tape = Tape(scene, spec.steps, 2)
    tape.forward(spec.dt)
    g, dp0, dv0 = tape.backward(dheight0=a, dheight1=b, dpos=c, dvel=d)
    tape.close()
    scene.close()
    names = SCALARS if couple else SCALARS[:-1]
    for name in names:
        fd = _fd_scalar(oracle, spec, spec.steps, name, loss)
        _say(f"{name}: adjoint {g[name]:.8e} fd {fd:.8e}")
        assert _rel(g[name], fd) < 2e-4, name
    # A 1e-6 step moves some of 4 096 particles across texel edges of the
    # 256^2 bilinear sampler, kinks the derivative does not see (the central
    # difference there is off by 8e-4; at 1e-7 and 1e-8 it has converged to
    # 2e-9): the wide case steps 1e-8.
    scale = 0.01 if wave >= 256 else 1.0
    for which, grad in (("pos", dp0), ("vel", dv0)):
        dirn = rng.normal(size=(n, 3))
        fd = _fd_direction(oracle, spec, spec.steps, which, dirn, loss, scale)
        ad = float((grad * dirn).sum())
        _say(f"d{which}0 . dir: adjoint {ad:.8e} fd {fd:.8e}")
        assert _rel(ad, fd) < 2e-4, which


def _reduction(cwa, oracle):
    from coupledwateranimation_amd import configs

    spec = configs.dam_break_diff(n=4096, wave=64, dt=2e-3, steps=32)
    # target heightfield: the forward model with each parameter scaled by
    # U[0.8, 1.2] (seed 3)
    f = np.random.default_rng(3).uniform(0.8, 1.2, 6)
    pt = spec.params.copy()
    pt.visc *= f[0]
    pt.gas_const *= f[1]
    pt.rest_rho *= f[2]
    pt.lambda_x *= f[3] ** 2  # wave speed c -> lambda ~ c^2
    pt.lambda_z *= f[3] ** 2
    pt.atten = min(pt.atten * f[4], 0.9999)
    pt.couple_strength *= f[5]
    target = _oracle_final(oracle, spec, pt, spec.steps)["height0"]
    return spec, target


def test_reduction_gradients_vs_fd(gpu, cwa, oracle):
    """4 096 particles, 64^2 texels, 32 steps, checkpoint every 8: MSE
    against a target from perturbed parameters.  Six scalars (lambda_x and
    lambda_z carry the wave speed) and velocity directions within 1e-3 of
    FP64 central differences."""
    from coupledwateranimation_amd.gradient import Tape

    spec, target = _reduction(cwa, oracle)
    N = target.size

    def loss(st):
        return float(((st["height0"] - target) ** 2).mean())

    scene = spec.make_scene()
    tape = Tape(scene, spec.steps, 8)
    tape.forward(spec.dt)
    h = scene.readback()["height0"].astype(np.float64)
    g, _, dv0 = tape.backward(dheight0=2.0 * (h - target) / N)
    tape.close()
    scene.close()
    for name in SCALARS:
        fd = _fd_scalar(oracle, spec, spec.steps, name, loss)
        _say(f"{name}: adjoint {g[name]:.8e} fd {fd:.8e} rel {_rel(g[name], fd):.2e}")
        assert _rel(g[name], fd) < 1e-3, name
    # initial velocities: along sign(gradient) (well conditioned) and along the
    # five largest components one at a time
    d = np.sign(dv0)
    fd = _fd_direction(oracle, spec, spec.steps, "vel", d, loss, 1.0)
    ad = float((dv0 * d).sum())
    _say(f"dv0 . sign: adjoint {ad:.8e} fd {fd:.8e}")
    assert _rel(ad, fd) < 1e-3
    flat = np.argsort(-np.abs(dv0).ravel())[:5]
    for idx in flat:
        e = np.zeros(dv0.size)
        e[idx] = 1.0
        fd = _fd_direction(oracle, spec, spec.steps, "vel", e.reshape(-1, 3), loss, 1.0)
        ad = float(dv0.ravel()[idx])
        _say(f"dv0[{idx}]: adjoint {ad:.8e} fd {fd:.8e}")
        assert _rel(ad, fd) < 1e-3


def test_checkpoint_interval_and_autograd(gpu, cwa, oracle):
    """Gradients do not depend on the checkpoint interval (up to the order
    of the fp32 atomics in the heightfield adjoint), and the
    torch.autograd.Function returns the same values."""
    import torch

    from coupledwateranimation_amd.gradient import Rollout, coupled_rollout

    spec, target = _reduction(cwa, oracle)
    spec.steps = 12
    res = []
    for every in (1, 5, 12):
        ro = Rollout(spec, 12, every)
        theta = torch.tensor(ro.theta0(1.0), dtype=torch.float64, requires_grad=True)
        v0 = torch.zeros((spec.n, 3), dtype=torch.float32, requires_grad=True)
        hgt, _ = coupled_rollout(theta, v0, ro)
        loss = ((hgt.double() - torch.from_numpy(target)) ** 2).mean()
        loss.backward()
        res.append((theta.grad.numpy().copy(), v0.grad.numpy().copy()))
    for t, v in res[1:]:
        np.testing.assert_allclose(t, res[0][0], rtol=1e-4)
        np.testing.assert_allclose(v, res[0][1], rtol=1e-3, atol=1e-4 * np.abs(res[0][1]).max())


def test_tape_and_fit_refuse_a_changed_scene(gpu, cwa):
    """A tape or fit sized for a scene's layout refuses to run (CWA_ESTATE,
    before any launch) once the scene was reinitialised or its grid, the
    sign of its coupling band or its wave mode changed; a change of plain
    scalars (a fit's business) keeps both usable."""
    from coupledwateranimation_amd import configs, fit
    from coupledwateranimation_amd.fit import FitBatch
    from coupledwateranimation_amd.gradient import Tape

    spec = configs.dam_break_diff(n=2048, wave=32, steps=4)
    sc = spec.make_scene()
    tape = Tape(sc, 4, 2)
    q = spec.params.copy()
    q.visc *= 1.5
    sc.set_params(q)                       # scalars only: still valid
    tape.forward(spec.dt)
    tape.backward(dheight0=np.ones((32, 32), np.float32))
    for change in ("grid", "band", "wake", "reinit"):
        r = spec.params.copy()
        if change == "grid":
            configs.grid_for_box(r, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 1.5 * r.h)
        elif change == "band":
            r.couple_band = 0.0
        elif change == "wake":
            r.wave_type = 0.5
        if change == "reinit":
            sc.reinit(spec.pos, spec.vel, None, spec.wave_size, spec.height0, spec.height1)
        else:
            sc.set_params(r)
        with pytest.raises(cwa.CwaError) as e:
            tape.forward(spec.dt)
        assert e.value.code == -4, change
        with pytest.raises(cwa.CwaError) as e:
            tape.backward(dheight0=np.ones((32, 32), np.float32))
        assert e.value.code == -4, change
        tape.close()
        sc.set_params(spec.params)
        tape = Tape(sc, 4, 2)             # a new tape on the new layout works
        tape.forward(spec.dt)
    tape.close()
    sc.close()

    specs, targets, _ = fit.fit_workload(nscenes=2, seed=5, steps=4, dt=5e-4,
                                         n_range=(1024, 2048), side_range=(32, 40))
    scenes = [s.make_scene() for s in specs]
    fb = FitBatch(scenes, targets, 4, 2, specs[0].dt)
    fb.run()
    r = specs[1].params.copy()
    r.couple_band = 0.0
    scenes[1].set_params(r)
    with pytest.raises(cwa.CwaError) as e:
        fb.run()
    assert e.value.code == -4
    fb.close()
    for s in scenes:
        s.close()


def test_pair_adjoint_lane_forms_agree(gpu, cwa, oracle):
    """The two forms of the density pass and the pair adjoint, four lanes
    per particle (scenes up to 2^18 particles) and one lane (larger scenes),
    forced on one small scene: a 2 048-particle dam break beside a
    compressed block (spacing h/2.8, up to 27 particles per cell, 90
    neighbours), whose long columns send particles through the re-search
    by distance in both forms.  Seeds on every output of 3 steps: parameter
    gradients and the initial-state gradients agree to 1e-4 (the forms sum
    in different orders)."""
    from coupledwateranimation_amd import configs
    from coupledwateranimation_amd.gradient import Tape
    from coupledwateranimation_amd.scene import Scene, set_lane_split_limit

    spec = configs.dam_break_diff(n=2048, wave=32, dt=1e-3, steps=3)
    p = spec.params
    sp = p.h / 2.8
    block = configs.lattice(12 ** 3, (0.55, 0.1, 0.55),
                            (0.55 + 12 * sp, 0.1 + 12 * sp, 0.55 + 12 * sp), sp, 0.5, 0.1,
                            (2, 0, 1), 11)
    pos = np.concatenate([spec.pos, block])
    n = pos.shape[0]
    rng = np.random.default_rng(7)
    vel = np.zeros_like(pos)
    vel[:, :3] = rng.normal(scale=0.2, size=(n, 3))
    g = oracle.grid_build(p, pos)
    assert p.grid_ydiv == 1 and g["count"].max() * 3 > 64  # some column past 64 candidates
    W, H = spec.wave_size
    a, b = rng.normal(size=(H, W)), rng.normal(size=(H, W))
    c, d = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))

    def run(limit):
        prev = set_lane_split_limit(limit)
        try:
            sc = Scene(p, pos, vel, None, (W, H), spec.height0, spec.height1)
            tape = Tape(sc, spec.steps, 2)
            tape.forward(spec.dt)
            out = tape.backward(dheight0=a, dheight1=b, dpos=c, dvel=d)
            tape.close()
            sc.close()
        finally:
            set_lane_split_limit(prev)
        return out

    g4, dp4, dv4 = run(-1)
    g1, dp1, dv1 = run(0)
    for name in SCALARS:
        _say(f"{name}: four lanes {g4[name]:.8e} one lane {g1[name]:.8e}")
        assert _rel(g1[name], g4[name]) < 1e-4, name
    for which, x1, x4 in (("pos", dp1, dp4), ("vel", dv1, dv4)):
        e = np.abs(x1 - x4).max() / np.abs(x4).max()
        _say(f"d{which}0: max difference {e:.3e} of the largest")
        assert e < 1e-4, which
