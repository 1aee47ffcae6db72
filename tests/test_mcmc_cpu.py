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
