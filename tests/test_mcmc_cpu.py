"""CPU checks of the sampler in lcurve_rs.mcmc on analytic targets: the
stretch-move ensemble samples a known Gaussian, delayed acceptance with a
deliberately wrong surrogate still targets the exact posterior (the second
stage corrects it), the exact-evaluation fallback engages when the
surrogate is far off, and the autocorrelation time of an AR(1) process
matches (1 + rho) / (1 - rho)."""
import numpy as np
import pytest

from lcurve_rs import mcmc

CENTER = np.array([0.5, 80.0, 0.015, 15000.0, 3500.0])
SCALE = np.array([0.02, 0.5, 0.0005, 300.0, 40.0])
CORR = np.array([[1.0, 0.6, 0.0, 0.0, 0.0],
                 [0.6, 1.0, 0.3, 0.0, 0.0],
                 [0.0, 0.3, 1.0, 0.0, 0.0],
                 [0.0, 0.0, 0.0, 1.0, -0.4],
                 [0.0, 0.0, 0.0, -0.4, 1.0]])
COV = CORR * np.outer(SCALE, SCALE)
PREC = np.linalg.inv(COV)


class GaussTarget:
    def __init__(self, bias=None, surr_scale=1.0):
        self.n_exact = 0
        self.bias = bias
        self.surr_scale = surr_scale

    def _lp(self, p):
        d = np.atleast_2d(p) - CENTER
        return -0.5 * np.einsum("ij,jk,ik->i", d, PREC, d)

    def logp_exact(self, p):
        p = np.atleast_2d(p)
        self.n_exact += len(p)
        return self._lp(p)

    def logp_surrogate(self, surr, p):
        p = np.atleast_2d(p)
        lp = self._lp(p) * self.surr_scale
        if self.bias is not None:
            lp = lp + ((p - CENTER) / SCALE) @ self.bias
        return lp


def _p0(nw, seed):
    rng = np.random.default_rng(seed)
    return CENTER + 0.1 * SCALE * rng.standard_normal((nw, 5))


def _check_posterior(chain, tol_mean=0.15, tol_std=0.15):
    c = chain[len(chain) // 4:].reshape(-1, 5)
    z = (c.mean(0) - CENTER) / SCALE
    assert np.all(np.abs(z) < tol_mean), z
    assert np.all(np.abs(c.std(0) / SCALE - 1) < tol_std), c.std(0) / SCALE


def test_plain_ensemble_samples_gaussian():
    r = mcmc.ensemble_mcmc(GaussTarget(), _p0(32, 0), 3000, seed=1)
    _check_posterior(r["chain"])
    assert r["n_exact"] == 32 + 3000 * 32


def test_delayed_acceptance_corrects_a_wrong_surrogate():
    """Surrogate 30% too wide and tilted: DA still samples the exact target,
    with fewer exact evaluations than proposals."""
    t = GaussTarget(bias=np.array([0.8, -0.5, 0.3, 0.0, 0.4]), surr_scale=0.7)
    r = mcmc.ensemble_mcmc(t, _p0(32, 2), 3000, seed=3, surrogate=object(), surr_tol=1e9)
    _check_posterior(r["chain"])
    assert r["n_exact"] < 0.7 * r["proposals"]
    assert r["fallback_at"] is None


def test_surrogate_transition_with_inner_steps():
    """Five surrogate-only moves per exact check (surrogate too wide and
    tilted): still the exact posterior, one exact evaluation per moved
    walker and sweep."""
    t = GaussTarget(bias=np.array([0.3, -0.2, 0.1, 0.0, 0.2]), surr_scale=0.85)
    r = mcmc.ensemble_mcmc(t, _p0(32, 6), 3000, seed=7, surrogate=object(), surr_tol=1e9, inner_steps=5)
    _check_posterior(r["chain"])
    assert r["n_exact"] <= 32 + 3000 * 32


def test_fallback_when_the_surrogate_is_far_off():
    """A surrogate whose error varies by ~50 in log-likelihood across the
    ensemble: the sampler switches to exact steps and stays correct."""
    t = GaussTarget(bias=np.array([50.0, 0, 0, 0, 0]))
    r = mcmc.ensemble_mcmc(t, _p0(32, 4), 3000, seed=5, surrogate=object(), surr_tol=5.0)
    assert r["fallback_at"] is not None and r["fallback_at"] < 10
    _check_posterior(r["chain"])


def test_autocorr_time_of_ar1():
    rng = np.random.default_rng(7)
    rho, n, nw = 0.8, 20000, 8
    x = np.zeros((n, nw, 1))
    e = rng.standard_normal((n, nw))
    for i in range(1, n):
        x[i, :, 0] = rho * x[i - 1, :, 0] + e[i]
    tau = mcmc.autocorr_time(x)[0]
    assert tau == pytest.approx((1 + rho) / (1 - rho), rel=0.1)


def test_autoscaled_chisq_recovers_exact_multiple():
    rng = np.random.default_rng(8)
    f = rng.uniform(1, 2, (3, 50))
    y = 2.5 * f[1]
    chi = mcmc.autoscaled_chisq(f, y, 0.01 * np.ones(50), np.ones(50))
    assert chi[1] == pytest.approx(0.0, abs=1e-18)
    assert chi[0] > 0 and chi[2] > 0


def test_box_mapping_round_trip():
    p = mcmc.sample_box(100, 3)
    assert np.all(mcmc.in_box(p))
    assert np.allclose(mcmc.from_unit(mcmc.to_unit(p)), p)
    assert np.all(np.abs(mcmc.to_unit(p)) <= 1.0)


def test_surrogate_curve_interpolates_periodically():
    s = mcmc.Surrogate.__new__(mcmc.Surrogate)
    s.nbins = 8
    s.times = None
    s.phases = mcmc.bin_phases(8)
    lg = np.log(np.arange(1, 9, dtype=float))[None, :]
    s.predict = lambda p: lg
    ph = np.array([0.0625, 0.125, 0.9375, 1.0, 1.0625 + 3.0])
    c = s.curve(np.zeros((1, 5)), ph)[0]
    # bin centres at (j + 0.5) / 8; phase 0 and 1 lie halfway between the last and first bins
    assert c[0] == pytest.approx(1.0)
    assert c[1] == pytest.approx(np.exp(0.5 * (np.log(1) + np.log(2))))
    assert c[2] == pytest.approx(8.0)
    assert c[3] == pytest.approx(np.exp(0.5 * (np.log(8) + np.log(1))))
    assert c[4] == pytest.approx(1.0)


def test_eclipse_phase_grid():
    ph = mcmc.eclipse_phases(256, 192, 0.06)
    assert len(ph) == 256 and np.all(np.diff(ph) > 0) and ph[0] >= 0 and ph[-1] < 1
    near = np.minimum(ph, 1 - ph) < 0.06
    assert near.sum() == 192


def test_time_output_surrogate_serves_its_own_times_only():
    s = mcmc.Surrogate.__new__(mcmc.Surrogate)
    s.times = np.array([-0.1, 0.0, 0.3, 1.1])
    s.nbins = 4
    lg = np.log(np.array([[1.0, 2.0, 3.0, 4.0]]))
    s.predict = lambda p: lg
    assert np.allclose(s.curve(np.zeros((1, 5)), s.times.copy()), [[1.0, 2.0, 3.0, 4.0]])
    with pytest.raises(ValueError):
        s.curve(np.zeros((1, 5)), s.times[:3])


def test_box_frame_is_the_unit_box():
    f = mcmc.Frame.from_box(mcmc.BOX)
    p = mcmc.sample_box(200, 4)
    assert np.allclose(f.to_u(p), mcmc.to_unit(p), rtol=0, atol=1e-12)
    assert np.allclose(f.from_u(f.to_u(p)), p)
    assert np.all(f.contains(p))
    out = p.copy()
    out[:, 1] = 95.0
    assert not np.any(f.contains(out))
    q = f.sample(300, 5)
    assert np.all(mcmc.in_box(q))


def test_local_frame_whitens_correlated_samples():
    rng = np.random.default_rng(9)
    x = rng.multivariate_normal(CENTER, COV, size=20000)
    f = mcmc.Frame.local(x, radius=5.0)
    u = f.to_u(x)
    # Unit covariance in units of the radius, centred on the samples' mean.
    assert np.allclose(np.cov(u, rowvar=False) * 25.0, np.eye(5), atol=1e-10)
    assert np.allclose(u.mean(0), 0.0, atol=1e-12)
    assert np.allclose(f.center, x.mean(0))
    # The ball holds all but the far tail: P(chi2_5 > 25) = 1.4e-4.
    assert 0.9995 < f.contains(x).mean() <= 1.0
    # Mahalanobis distance under the true covariance tracks |u| * radius.
    d_true = np.sqrt(np.einsum("ij,jk,ik->i", x - CENTER, PREC, x - CENTER))
    assert np.corrcoef(d_true, 5.0 * np.linalg.norm(u, axis=1))[0, 1] > 0.999


def test_ball_frame_sampling():
    rng = np.random.default_rng(10)
    f = mcmc.Frame.local(rng.multivariate_normal(CENTER, COV, size=5000), radius=6.0)
    p = f.sample(4001, 11)
    assert p.shape == (4001, 5)
    r = np.linalg.norm(f.to_u(p), axis=1)
    assert np.all(r <= 1.0 + 1e-12)
    # Half from the core (a normal of sd 1/3 of the radius cut to the ball:
    # 9 r^2 is chi-squared with 5 degrees of freedom), the rest uniform in
    # the ball (P(r < x) = x^5).
    from scipy.stats import chi2
    cut = chi2.cdf(9.0, 5)
    for x in (0.3, 0.5, 0.7, 0.9):
        expect = 0.5 * chi2.cdf(9.0 * x * x, 5) / cut + 0.5 * x ** 5
        assert abs(np.mean(r < x) - expect) < 0.03, (x, np.mean(r < x), expect)
    assert np.array_equal(p, f.sample(4001, 11))


def test_polynomial_features():
    u = np.random.default_rng(12).uniform(-1, 1, (7, 5))
    from math import comb
    for deg in range(5):
        assert mcmc.poly_features(u, deg).shape == (7, comb(5 + deg, deg))
    assert np.array_equal(mcmc.poly_features(u, 2), mcmc.quad_features(u))
    F3 = mcmc.poly_features(u, 3)
    # Columns: 1, u0..u4, then u_i u_j (i <= j), then u_i u_j u_k.
    assert np.allclose(F3[:, 0], 1.0)
    assert np.allclose(F3[:, 1:6], u)
    assert np.allclose(F3[:, 6], u[:, 0] ** 2)
    assert np.allclose(F3[:, 21], u[:, 0] ** 3)
    assert np.allclose(F3[:, -1], u[:, 4] ** 3)
    # A cubic is fitted exactly.
    x = np.random.default_rng(13).uniform(-1, 1, (200, 5))
    y = 1.0 + x[:, 0] * x[:, 1] * x[:, 2] - 2.0 * x[:, 4] ** 3 + 0.5 * x[:, 3]
    c = np.linalg.lstsq(mcmc.poly_features(x, 3), y, rcond=None)[0]
    assert np.allclose(mcmc.poly_features(u, 3) @ c, 1.0 + u[:, 0] * u[:, 1] * u[:, 2] - 2.0 * u[:, 4] ** 3
                       + 0.5 * u[:, 3])


class _FakeTarget(mcmc.Target):
    """A Target with observed data but no exact model (no GPU needed)."""

    def __init__(self, times, y, ferr):
        self.times, self.y, self.ferr, self.w = times, y, ferr, np.ones_like(y)
        self._lik = None


class _LinearSurrogate:
    """flux = exp(0.01 * u . a) at every time: a stand-in with a frame."""
    times = None
    version = 0

    def __init__(self, frame, tail):
        self.frame, self.tail = frame, tail
        self.a = np.linspace(0.5, 1.5, 5)

    def contains(self, p):
        return self.frame.contains(p)

    def curve(self, p, t):
        u = self.frame.to_u(p)
        return np.exp(0.01 * (u @ self.a)[:, None] * np.linspace(-1.0, 1.0, len(t))[None, :])


def test_surrogate_gaussian_tail_outside_the_ball():
    rng = np.random.default_rng(14)
    pilot = rng.multivariate_normal(CENTER, COV, size=4000)
    frame = mcmc.Frame.local(pilot, radius=3.0)
    t = np.linspace(0.0, 1.0, 40)
    y = 1.0 + 0.01 * rng.standard_normal(40)
    tg = _FakeTarget(t, y, 0.01 * np.ones(40))
    p = np.concatenate([frame.sample(50, 15), rng.multivariate_normal(CENTER, 25.0 * COV, size=50)])
    ins = frame.contains(p)
    assert 10 < ins.sum() < 100
    expect_in = -0.5 * mcmc.autoscaled_chisq(_LinearSurrogate(frame, False).curve(p[ins], t), y, tg.ferr, tg.w)
    plain = tg.logp_surrogate(_LinearSurrogate(frame, False), p)
    assert np.allclose(plain[ins], expect_in, rtol=1e-13) and np.all(plain[~ins] == -np.inf)
    s = _LinearSurrogate(frame, True)
    tailed = tg.logp_surrogate(s, p)
    assert np.array_equal(tailed[ins], plain[ins])
    l0 = -0.5 * mcmc.autoscaled_chisq(s.curve(frame.center[None], t), y, tg.ferr, tg.w)[0]
    # Squared Mahalanobis distance under the pilot samples' covariance.
    cov = np.cov(pilot, rowvar=False)
    d2 = np.einsum("ij,jk,ik->i", p[~ins] - frame.center, np.linalg.inv(cov), p[~ins] - frame.center)
    assert np.allclose(tailed[~ins], l0 - 0.5 * d2, rtol=1e-12)
    # Past the ball the tail is below its value at the boundary.
    assert np.all(tailed[~ins] < l0 - 0.5 * 9.0 + 1e-9)
    with pytest.raises(ValueError):
        mcmc.Surrogate.__init__(mcmc.Surrogate.__new__(mcmc.Surrogate), tail=True)


def test_autoscaled_chisq_is_re_scale_with_masked_points(golden_model_text):
    """mcmc.autoscaled_chisq (the host twin of the GPU surrogate likelihood)
    follows re_scale (orchestration.rs:361-388) as the oracle restates it:
    points with weight <= 0 take no part even with ferr = 0 or a negative
    weight, and with no usable point the chi-squared is 0."""
    import pyoracle
    from lcurve_rs import workloads as W
    t = W.linspace_times(-0.2, 1.2, 90)
    base, _ = pyoracle.light_curve(golden_model_text, t)
    rng = np.random.default_rng(5)
    y = 1.7 * base * (1 + 0.01 * rng.standard_normal(len(t)))
    ferr = 0.01 * np.abs(y)
    w = np.ones(len(t))
    w[::5] = 0.0
    ferr[::10] = 0.0          # zero error on points of zero weight
    w[3] = -1.0               # a negative weight is excluded too
    _, sc = pyoracle.light_curve(golden_model_text, t, flux=y, ferr=ferr, weight=w, scale=True, rdata=True)
    got = mcmc.autoscaled_chisq(base[None, :], y, ferr, w)[0]
    assert np.isfinite(got) and got == pytest.approx(sc["chisq"], rel=1e-12)
    assert mcmc.autoscaled_chisq(base[None, :], y, ferr, np.zeros(len(t)))[0] == 0.0
    assert mcmc.autoscaled_chisq(np.zeros((1, len(t))), y, ferr, w)[0] == 0.0
