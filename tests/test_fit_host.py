"""Host logic of the batched fit's descent (fit.FitRun.descend), with the
graph evaluation replaced by closed-form losses: per-scene L-BFGS in the
scaled fit coordinates, reject-and-halve step control, bounds.  No GPU."""
import numpy as np
import pytest

from coupledwateranimation_amd import fit


class _Stub(fit.FitRun):
    """A FitRun whose evaluate() is a separable quadratic per scene,
    loss_k(x) = sum_i w_ki (x_i - t_ki)^2, optionally non-finite beyond a
    radius of the start (to drive the reject path)."""

    def __init__(self, targets, weights, blowup=None):
        ns = len(targets)
        self.scenes = [None] * ns
        self.scene_bytes = [0] * ns
        self.theta0 = np.tile(np.array([1.0, 1.0, 1.0, 1.0, 0.999, 1.0]), (ns, 1))
        self.t = np.asarray(targets, float)
        self.w = np.asarray(weights, float)
        self.blowup = blowup
        self.calls = 0

        class _F:
            def tape_bytes(self, s):
                return 0

        self.fit = _F()

    def evaluate(self, x):
        self.calls += 1
        d = x - self.t
        loss = (self.w * d * d).sum(axis=1)
        g = 2.0 * self.w * d
        if self.blowup is not None:
            far = np.abs(x - self.x0()).max(axis=1) > self.blowup
            loss = np.where(far, np.inf, loss)
        return loss, g


def _targets(ns, seed):
    rng = np.random.default_rng(seed)
    st = _Stub(np.zeros((ns, 6)), np.ones((ns, 6)))
    x0 = st.x0()
    t = x0 + rng.uniform(-0.15, 0.15, (ns, 6))
    w = 10.0 ** rng.uniform(-1, 1, (ns, 6))       # condition number up to 100
    return t, w


def test_lbfgs_reaches_the_minimum_of_each_scene():
    t, w = _targets(5, 1)
    st = _Stub(t, w)
    r = st.descend(iters=60)
    assert st.calls == 61                            # one evaluation per iteration
    assert np.all(np.diff(r.accepted, axis=0) <= 0)  # the accepted loss never rises
    assert np.all(r.best_loss <= 1e-8 * r.losses[0])
    assert np.allclose(r.x[-1], t, atol=1e-4)


def test_rejected_proposals_halve_the_step():
    t, w = _targets(3, 2)
    st = _Stub(t, w, blowup=0.05)                    # every step longer than 0.05 fails
    r = st.descend(iters=40)
    assert np.all(r.rejected > 0)
    assert np.all(np.isfinite(r.accepted[-1]))
    assert np.all(r.accepted[-1] < r.accepted[0])   # progress inside the trust region
    assert np.all(np.abs(r.x[-1] - st.x0()) <= 0.05 + 1e-12)


def test_bounds_clip_the_proposals():
    ns = 2
    st0 = _Stub(np.zeros((ns, 6)), np.ones((ns, 6)))
    t = st0.x0() + 5.0                               # minimum far outside X_HI
    st = _Stub(t, np.ones((ns, 6)))
    r = st.descend(iters=30)
    assert np.all(r.x[-1] <= fit.X_HI + 1e-12) and np.all(r.x[-1] >= fit.X_LO - 1e-12)
    # the multiplicative scalars end on their upper bound
    mult = [i for i in range(6) if i != fit.ATTEN]
    assert np.allclose(r.x[-1][:, mult], fit.X_HI[mult], atol=1e-6)
