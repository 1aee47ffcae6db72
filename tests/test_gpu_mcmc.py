"""GPU checks of the surrogate pipeline (lcurve_rs.mcmc): training curves
from the batched exact path equal single light curves, the surrogate
(quadratic baseline + BF16 MLP residual) learns a local box, and the
delayed-acceptance sampler runs on the exact GPU likelihood with fewer or
equal exact evaluations than proposals."""
import numpy as np
import pytest

from lcurve_rs import workloads as W

pytestmark = pytest.mark.gpu

NLAT = (10, 10, 20, 20)


@pytest.fixture(scope="module")
def mc(lcurve_rs):
    from lcurve_rs import mcmc
    return mcmc


def test_training_curves_equal_light_curves(lcurve_rs, mc):
    base = mc.surrogate_base_text(NLAT)
    p = mc.sample_box(3, 11)
    ph = mc.eclipse_phases(128, 64, 0.06)
    lg = mc.make_training_set(base, p, phases=ph)
    for i in range(3):
        m = lcurve_rs.Model.from_text(W.model_text(dict(zip(mc.PARAM_NAMES, p[i])), base))
        r = m.light_curve(times=ph, expose=0.0, ndiv=1)
        assert np.allclose(lg[i], np.log(r.calc / r.wdwarf), rtol=0, atol=1e-12)


def test_surrogate_learns_a_local_box(mc):
    base = mc.surrogate_base_text(NLAT)
    center = np.array([0.5, 82.0, 0.015, 15000.0, 3500.0])
    box = mc.local_box(center, center * 0.005, 1.0)
    p = mc.sample_box(3000, 5, box)
    lg = mc.make_training_set(base, p, 128)
    s = mc.Surrogate(128, hidden=(256, 256), seed=2, box=box)
    s.fit(p[500:], lg[500:], epochs=150, batch=512)
    err = s.predict(p[:500]) - lg[:500]
    quad = mc.quad_features(mc.to_unit(p[:500], box)) @ s.coef - lg[:500]
    assert np.sqrt(np.mean(err ** 2)) < 2e-3
    assert np.sqrt(np.mean(err ** 2)) <= np.sqrt(np.mean(quad ** 2)) * 1.05


def test_delayed_acceptance_runs_on_the_exact_likelihood(mc):
    base = mc.surrogate_base_text(NLAT)
    truth = np.array([0.5, 82.0, 0.015, 15000.0, 3500.0])
    tg = mc.Target(base, truth, np.linspace(-0.2, 1.2, 300), 0.02, seed=1)
    box = mc.local_box(truth, truth * 0.02, 1.0)
    p = mc.sample_box(2000, 9, box)
    s = mc.Surrogate(128, hidden=(256, 256), seed=3, box=box)
    s.fit(p, mc.make_training_set(base, p, 128), epochs=100, batch=512)
    rng = np.random.default_rng(0)
    p0 = truth * (1 + 1e-3 * rng.standard_normal((16, 5)))
    r = mc.ensemble_mcmc(tg, p0, 20, seed=1, surrogate=s, inner_steps=2, surr_tol=1e9)
    assert r["n_exact"] <= 16 + 20 * 16
    assert np.all(mc.in_box(r["chain"].reshape(-1, 5), box))
    assert 0 < r["accepted"] <= r["proposals"]
