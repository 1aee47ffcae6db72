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
    """In a box of +-5% about the centre the quadratic baseline misses
    structure the network learns: the network cuts the quadratic's held-out
    error at least fourfold (measured: 8.8e-3 -> 9.5e-4). In much smaller
    boxes what the polynomial leaves is the surface grids' discretisation
    noise, which no smooth regressor can learn."""
    base = mc.surrogate_base_text(NLAT)
    center = np.array([0.5, 82.0, 0.015, 15000.0, 3500.0])
    box = mc.local_box(center, center * 0.05, 1.0)
    p = mc.sample_box(3000, 5, box)
    lg = mc.make_training_set(base, p, 128)
    s = mc.Surrogate(128, hidden=(256, 256), seed=2, box=box)
    s.fit(p[500:], lg[500:], epochs=400, batch=512)
    err = s.predict(p[:500]) - lg[:500]
    quad = mc.quad_features(mc.to_unit(p[:500], box)) @ s.coef - lg[:500]
    rms_net, rms_quad = np.sqrt(np.mean(err ** 2)), np.sqrt(np.mean(quad ** 2))
    assert rms_net < 1.5e-3
    # the network's gain over the method without it: the quadratic baseline
    # alone leaves at least four times the held-out error
    assert rms_net <= rms_quad / 4.0, (rms_net, rms_quad)


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


def test_time_output_surrogate_in_a_local_frame(lcurve_rs, mc):
    """A surrogate of a smeared target's own curve: training curves at the
    data times with exposure smearing equal single light curves, the cubic
    baseline in whitened coordinates plus the network beats the baseline
    alone on held-out curves, and the surrogate's likelihood is finite
    inside its ball and -inf outside."""
    base = mc.surrogate_base_text(NLAT)
    truth = np.array([0.5, 82.0, 0.015, 15000.0, 3500.0])
    times = np.linspace(-0.2, 1.2, 300)
    expose, ndiv = 0.002, 3
    rng = np.random.default_rng(4)
    sd = np.array([0.004, 0.2, 0.0002, 150.0, 15.0])
    corr = np.eye(5)
    corr[0, 1] = corr[1, 0] = -0.8
    pilot = rng.multivariate_normal(truth, corr * np.outer(sd, sd), size=400)
    frame = mc.Frame.local(pilot, radius=20.0)
    p = frame.sample(2500, 6)
    lg = mc.make_training_set(base, p, times=times, expose=expose, ndiv=ndiv)
    m = lcurve_rs.Model.from_text(W.model_text(dict(zip(mc.PARAM_NAMES, p[0])), base))
    r = m.light_curve(times=times, expose=expose, ndiv=ndiv)
    assert np.allclose(lg[0], np.log(r.calc / r.wdwarf), rtol=0, atol=1e-12)
    s = mc.Surrogate(hidden=(256, 256), seed=2, frame=frame, times=times, degree=3, tail=False)
    assert s.nout == 320 and s.nbins == 300
    s.fit(p[400:], lg[400:], epochs=400, batch=512, lr=3e-3)
    err = s.predict(p[:400]) - lg[:400]
    poly = s.baseline(p[:400]) - lg[:400]
    rms_net, rms_poly = np.sqrt(np.mean(err ** 2)), np.sqrt(np.mean(poly ** 2))
    assert rms_net < 1e-3
    # the network's gain: over a ball of 20 pilot deviations the cubic
    # baseline alone leaves at least 1.6 times the held-out error (measured:
    # 1.58e-3 -> 7.5e-4)
    assert rms_net <= rms_poly / 1.6, (rms_net, rms_poly)
    tg = mc.Target(base, truth, times, 0.02, seed=1, expose=expose, ndiv=ndiv)
    far = truth + 20.0 * sd
    ls = tg.logp_surrogate(s, np.stack([truth, far]))
    assert np.isfinite(ls[0]) and ls[1] == -np.inf
    le = tg.logp_exact(truth[None])
    assert abs(ls[0] - le[0]) < 5.0
    p0 = pilot[:16]
    res = mc.ensemble_mcmc(tg, p0, 10, seed=1, surrogate=s, inner_steps=4, surr_tol=1e9)
    assert np.all(frame.contains(res["chain"].reshape(-1, 5)))
    assert 0 < res["accepted"] <= res["proposals"]


def test_gpu_surrogate_likelihood_matches_host(mc):
    """The one-call GPU likelihood of a time-output surrogate (baseline +
    network + exp + autoscaled chi-squared on the device) equals the host's
    evaluation of the same surrogate, over more rows than one device chunk,
    with a row outside the domain, and follows a refit."""
    base = mc.surrogate_base_text(NLAT)
    truth = np.array([0.5, 82.0, 0.015, 15000.0, 3500.0])
    sd = np.array([0.004, 0.2, 0.0002, 150.0, 15.0])
    times = np.linspace(-0.2, 1.2, 500)
    expose, ndiv = 0.002, 3
    rng = np.random.default_rng(5)
    frame = mc.Frame.local(truth + rng.standard_normal((300, 5)) * sd, radius=5.0)
    p = frame.sample(640, 8)
    lg = mc.make_training_set(base, p, times=times, expose=expose, ndiv=ndiv)
    s = mc.Surrogate(hidden=(128,), seed=4, frame=frame, times=times, degree=3, tail=False)
    assert s.nout == 512
    s.fit(p, lg, epochs=5, batch=128)
    tg = mc.Target(base, truth, times, 0.01, seed=2, expose=expose, ndiv=ndiv)
    # masked points (re_scale drops weight <= 0), some with zero error
    tg.w[::7] = 0.0
    tg.ferr[::14] = 0.0
    q = np.concatenate([frame.sample(299, 9), (truth + 30.0 * sd)[None]])
    g = tg.logp_surrogate(s, q)
    h = tg.logp_surrogate_host(s, q)
    assert np.isinf(g[-1]) and np.isinf(h[-1])
    # the same masked points in the exact likelihood: finite, same scale
    assert np.all(np.isfinite(tg.logp_exact(q[:3])))
    assert np.all(np.isfinite(g[:-1])) and np.all(np.isfinite(h[:-1]))
    assert np.allclose(g[:-1], h[:-1], rtol=1e-9, atol=1e-7)
    assert np.all(np.isinf(tg.logp_surrogate(s, q[-1:])))
    s.fit(p, lg, epochs=1, batch=128)
    g2, h2 = tg.logp_surrogate(s, q[:5]), tg.logp_surrogate_host(s, q[:5])
    assert not np.allclose(g2, g[:5], rtol=1e-12, atol=0)
    assert np.allclose(g2, h2, rtol=1e-9, atol=1e-7)
