"""Host-side pieces of the gradient path, no GPU: the chain rule from the
library's parameter gradients to the six fit scalars (Rollout.chain_theta,
against finite differences of Rollout.params_for), and the fit's coordinate
maps (theta_of / x0, the scaled targets and the error measure)."""
import numpy as np
import pytest

from coupledwateranimation_amd import configs, fit
from coupledwateranimation_amd.gradient import Rollout

KEYS = ("visc", "gas_const", "rest_rho", "lambda_x", "lambda_z", "atten", "couple_strength")


@pytest.mark.parametrize("wave,c", [(64, 1.0), (96, 0.7)])
def test_chain_theta_matches_finite_differences(wave, c):
    spec = configs.dam_break_diff(n=512, wave=wave, seed=2, dt=3e-4, steps=8)
    ro = Rollout(spec, 8, 4)
    rng = np.random.default_rng(wave)
    g = {k: float(v) for k, v in zip(KEYS, rng.normal(size=len(KEYS)))}
    theta = np.array(ro.theta0(c), np.float64)

    def loss(th):
        # a loss linear in the library's parameters, with gradient g
        p = ro.params_for(th)
        return sum(g[k] * float(getattr(p, k)) for k in KEYS)

    got = ro.chain_theta(g, theta)
    for i in range(6):
        h = 1e-3 * max(abs(theta[i]), 1e-3)
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (loss(tp) - loss(tm)) / (2 * h)
        # the parameters are float32: central differences to ~1e-3 relative
        assert got[i] == pytest.approx(fd, rel=2e-3, abs=1e-6 * (1 + abs(fd))), (i, got[i], fd)


class _Coords(fit.FitRun):
    def __init__(self, theta0):
        self.theta0 = np.asarray(theta0, np.float64)
        self.scenes = [None] * len(self.theta0)


def test_fit_coordinates_round_trip_and_errors():
    th0 = np.array([[0.1, 3.0, 1000.0, 1.0, 0.999, 0.5], [0.2, 5.0, 900.0, 0.8, 0.995, 1.0]])
    run = _Coords(th0)
    x0 = run.x0()
    assert np.allclose(x0[:, fit.ATTEN], np.log(1.0 - th0[:, fit.ATTEN]))
    assert np.allclose(run.theta_of(x0), th0, rtol=1e-12)
    assert np.all(x0 >= fit.X_LO) and np.all(x0 <= fit.X_HI)
    # targets: the factors scale the scalars, and 1 - atten for the attenuation
    f = np.array([[1.1, 0.9, 1.05, 0.95, 1.2, 0.8], [0.85, 1.15, 1.0, 1.1, 0.9, 1.2]])
    truth = fit.true_theta(run, f)
    assert np.allclose(truth[:, :4], th0[:, :4] * f[:, :4])
    assert np.allclose(1.0 - truth[:, fit.ATTEN], (1.0 - th0[:, fit.ATTEN]) * f[:, fit.ATTEN])
    assert np.allclose(fit.theta_errors(truth, truth), 0.0)
    e = fit.theta_errors(th0, truth)
    assert np.allclose(e[:, :4], np.abs(1.0 / f[:, :4] - 1.0))
    assert np.allclose(e[:, fit.ATTEN], np.abs(1.0 / f[:, fit.ATTEN] - 1.0))
