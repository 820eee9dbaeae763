"""Batched heightfield fits (cwa_fit_*): one hipGraph holding every scene's
checkpointed forward, MSE seed and adjoint sweep.

Bars:
  loss of the graph = the host MSE of the eagerly stepped scene ... 1e-9 rel
  gradients of the graph = the single-tape eager path ............. 1e-4 rel
    (the heightfield adjoint accumulates with fp32 atomics, whose order
    differs between runs)
  parameters changed between replays take effect (device-resident
  parameter block) ................................................ 1e-4 rel
  a short descent lowers the loss of every scene and the identifiable
  scalars' error (the small batch and the 64-scene workload)
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SCALARS = ("visc", "gas_const", "rest_rho", "lambda_x", "lambda_z", "atten", "couple_strength")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _small(nscenes=3, steps=24):
    from coupledwateranimation_amd import fit

    return fit.fit_workload(nscenes=nscenes, seed=11, steps=steps, dt=5e-4,
                            n_range=(2048, 6000), side_range=(48, 96))


def _eager(spec, params, target, steps, every):
    from coupledwateranimation_amd.gradient import Tape
    from coupledwateranimation_amd.scene import Scene

    with Scene(params, spec.pos, spec.vel, None, spec.wave_size, spec.height0,
               spec.height1) as sc:
        tape = Tape(sc, steps, every)
        tape.forward(spec.dt)
        h = sc.readback()["height0"].astype(np.float64)
        t = target.astype(np.float64)
        loss = float(((h - t) ** 2).mean())
        g, _, _ = tape.backward(dheight0=2.0 * (h - t) / h.size)
        tape.close()
    return loss, g


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-30)


def test_fit_graph_matches_eager_tapes(gpu, cwa):
    from coupledwateranimation_amd.fit import FitBatch
    from coupledwateranimation_amd.scene import Scene

    specs, targets, _ = _small()
    scenes = [Scene(sp.params, sp.pos, sp.vel, None, sp.wave_size, sp.height0, sp.height1)
              for sp in specs]
    fb = FitBatch(scenes, targets, specs[0].steps, 8, specs[0].dt)
    losses, grads = fb.run()
    losses2, grads2 = fb.run()          # a second replay: same state, same answer
    for s, sp in enumerate(specs):
        _say(f"scene {s}: n={sp.n} wave={sp.wave_size} loss {losses[s]:.6e}")
        assert fb.tape_bytes(s) > 0
        loss_e, g_e = _eager(sp, sp.params, targets[s], sp.steps, 8)
        assert _rel(losses[s], loss_e) < 1e-9
        assert losses2[s] == losses[s]
        for k in SCALARS:
            assert _rel(grads[s][k], g_e[k]) < 1e-4, (s, k, grads[s][k], g_e[k])
            assert _rel(grads2[s][k], grads[s][k]) < 1e-4
    fb.close()
    for sc in scenes:
        sc.close()


def test_fit_graph_reads_updated_parameters(gpu, cwa):
    from coupledwateranimation_amd.fit import FitBatch
    from coupledwateranimation_amd.gradient import Rollout
    from coupledwateranimation_amd.scene import Scene

    specs, targets, _ = _small(nscenes=2)
    scenes = [Scene(sp.params, sp.pos, sp.vel, None, sp.wave_size, sp.height0, sp.height1)
              for sp in specs]
    fb = FitBatch(scenes, targets, specs[0].steps, 8, specs[0].dt)
    fb.run()
    new_params = []
    for sc, sp in zip(scenes, specs):
        ro = Rollout(sp, sp.steps, 8)
        th = np.array(ro.theta0(1.0)) * np.array([1.1, 0.9, 1.05, 1.1, 0.9999, 0.8])
        p = ro.params_for(th)
        sc.set_params(p)
        new_params.append(p)
    losses, grads = fb.run()
    for s, sp in enumerate(specs):
        loss_e, g_e = _eager(sp, new_params[s], targets[s], sp.steps, 8)
        _say(f"scene {s}: loss graph {losses[s]:.6e} eager {loss_e:.6e}")
        assert _rel(losses[s], loss_e) < 1e-9
        for k in SCALARS:
            assert _rel(grads[s][k], g_e[k]) < 1e-4, (s, k)
    fb.close()
    for sc in scenes:
        sc.close()


def test_descent_lowers_every_loss(gpu, cwa):
    """The accepted loss of every scene never rises (rejected proposals
    halve the step) and ends below the initial one, and the mean error of
    the identifiable scalars (wave speed, attenuation) falls."""
    from coupledwateranimation_amd import fit

    specs, targets, truth = _small(nscenes=3, steps=24)
    with fit.FitRun(specs, targets, ckpt_every=8) as run:
        res = run.descend(iters=12)
        truth_th = fit.true_theta(run, truth)
        e0 = fit.theta_errors(run.theta0, truth_th)
        e1 = fit.theta_errors(res.best_theta, truth_th)
    _say("losses", res.accepted[0], "->", res.accepted[-1], "rejected", res.rejected)
    _say("identifiable error", e0[:, [3, 4]].mean(), "->", e1[:, [3, 4]].mean())
    assert np.all(np.isfinite(res.accepted))
    assert np.all(np.diff(res.accepted, axis=0) <= 0.0)
    assert np.all(res.accepted[-1] < res.accepted[0])
    assert np.all(res.best_loss == res.accepted[-1])
    assert np.all(res.tape_bytes > 0) and np.all(res.scene_bytes > 0)
    assert e1[:, [3, 4]].mean() < e0[:, [3, 4]].mean()


def test_fit_workload_descent(gpu, cwa):
    """20 iterations of the descent on the 64-scene fit workload: every
    scene's accepted loss never rises and ends at most a tenth of its
    initial value, and the mean error of the wave speed and of the
    attenuation each fall below a fifth of their initial values."""
    from coupledwateranimation_amd import fit

    specs, targets, factors = fit.fit_workload(nscenes=64)
    with fit.FitRun(specs, targets) as run:
        res = run.descend(iters=20)
        truth = fit.true_theta(run, factors)
        e0 = fit.theta_errors(run.theta0, truth).mean(0)
        e1 = fit.theta_errors(res.best_theta, truth).mean(0)
    ratio = res.accepted[-1] / res.accepted[0]
    _say(f"64 scenes, 20 iterations: loss ratio median {np.median(ratio):.2e}, max "
         f"{ratio.max():.2e}; rejected median {np.median(res.rejected)}")
    _say("errors", {n: (round(a, 4), round(b, 4)) for n, a, b in zip(fit.THETA_NAMES, e0, e1)})
    assert np.all(np.isfinite(res.accepted))
    assert np.all(np.diff(res.accepted, axis=0) <= 0.0)
    assert np.all(ratio <= 0.1)
    assert e1[3] < 0.2 * e0[3] and e1[4] < 0.2 * e0[4]


def test_fit_workload_64_scenes(gpu, cwa, oracle):
    """The batched-fit workload itself: 64 scenes of 8k-128k particles on
    128^2-1024^2 heightfields, 128 steps checkpointed every 16, all in one
    graph.  For every scene the loss and the gradients are finite and the
    graph's loss equals the host MSE of the scene's final heightfield
    (1e-9 rel); for the smallest and the largest scene the gradients equal a
    single eager tape's (1e-4 rel); the smallest scene's 128-step forward is
    within 3x the envelope of FP32 oracle runs (the oracle and four runs from
    positions moved one ulp) of the FP64 oracle, per field."""
    from coupledwateranimation_amd import fit
    from coupledwateranimation_amd.fit import FitBatch
    from coupledwateranimation_amd.scene import Scene

    specs, targets, _ = fit.fit_workload(nscenes=64)
    ns = np.array([sp.n for sp in specs])
    _say(f"fit workload: {len(specs)} scenes, {ns.sum()} particles, steps {specs[0].steps}")
    scenes = [Scene(sp.params, sp.pos, sp.vel, None, sp.wave_size, sp.height0, sp.height1)
              for sp in specs]
    fb = FitBatch(scenes, targets, specs[0].steps, 16, specs[0].dt)
    losses, grads = fb.run()
    assert np.all(np.isfinite(losses)) and np.all(losses > 0)
    for s, sc in enumerate(scenes):
        assert all(np.isfinite(grads[s][k]) for k in SCALARS), s
        h = sc.readback()["height0"].astype(np.float64)
        mse = float(((h - targets[s].astype(np.float64)) ** 2).mean())
        assert _rel(losses[s], mse) < 1e-9, (s, losses[s], mse)
    small, large = int(np.argmin(ns)), int(np.argmax(ns))
    for s in (small, large):
        sp = specs[s]
        loss_e, g_e = _eager(sp, sp.params, targets[s], sp.steps, 16)
        _say(f"scene {s}: n={sp.n} wave={sp.wave_size} loss graph {losses[s]:.6e} "
             f"eager {loss_e:.6e}")
        assert _rel(losses[s], loss_e) < 1e-9
        for k in SCALARS:
            assert _rel(grads[s][k], g_e[k]) < 1e-4, (s, k, grads[s][k], g_e[k])
    st = scenes[small].readback()
    fb.close()
    for sc in scenes:
        sc.close()
    sp = specs[small]
    ref64 = oracle.OracleScene(sp.params, sp.pos, sp.vel, None, sp.wave_size, sp.height0,
                               sp.height1, double=True, threads=16)
    ref64.step(sp.dt, sp.steps)
    f64 = ref64.state()
    env = {}
    for k in range(5):
        pos = sp.pos.copy()
        if k:
            m = np.random.default_rng(k).random(pos[:, :3].shape) < 0.5
            pos[:, :3] = np.where(m, np.nextafter(pos[:, :3], np.float32(2)), pos[:, :3])
        o = oracle.OracleScene(sp.params, pos, sp.vel, None, sp.wave_size, sp.height0,
                               sp.height1, threads=16)
        o.step(sp.dt, sp.steps)
        f32 = o.state()
        o.close()
        for key, sl in (("pos", np.s_[:, :3]), ("vel", np.s_[:, :3]), ("rho", np.s_[:]),
                        ("height0", np.s_[:])):
            e = np.abs(f32[key][sl] - f64[key][sl]).max() / max(np.abs(f64[key][sl]).max(),
                                                                1e-30)
            env[key] = max(env.get(key, 0.0), e)
    ref64.close()
    for key, sl in (("pos", np.s_[:, :3]), ("vel", np.s_[:, :3]), ("rho", np.s_[:]),
                    ("height0", np.s_[:])):
        e = np.abs(st[key][sl] - f64[key][sl]).max() / max(np.abs(f64[key][sl]).max(), 1e-30)
        _say(f"smallest scene {key}: gpu-vs-fp64 {e:.3e}, fp32 envelope {env[key]:.3e}")
        assert e <= 3 * env[key] + 1e-6, key
