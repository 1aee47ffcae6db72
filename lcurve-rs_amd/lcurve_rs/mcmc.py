"""Surrogate-assisted MCMC over light-curve models.

* ``make_training_set`` — curves from the exact GPU method over a parameter
  box, folded to phase bins: log(flux / wdwarf) at the bin centres.
* ``Surrogate`` — the GPU MLP (lcurve_rs.surrogate.Mlp) trained on them,
  with the input / output normalisation and the phase interpolation that
  turns its bins into a curve at arbitrary phases.
* ``ensemble_mcmc`` — the affine-invariant stretch-move ensemble sampler
  (Goodman & Weare 2010) with exact likelihoods, or with delayed
  acceptance (Christen & Fox 2005) generalised to a surrogate transition
  (Liu 2001): each walker of a half-ensemble first makes ``inner_steps``
  stretch moves on the surrogate alone, the other half held fixed (a
  kernel reversible for the surrogate posterior), and the endpoint y is
  then accepted with probability min(1, pi(y) pi~(x) / (pi(x) pi~(y))),
  pi exact and pi~ surrogate, which makes the move reversible for the exact
  posterior. With inner_steps = 1 this is two-stage delayed acceptance; a
  walker whose inner moves were all rejected needs no exact evaluation.
  After every sweep the spread
  (standard deviation over the walkers) of surrogate minus exact
  log-likelihood at the current states — all known exactly — is checked;
  a constant offset cancels in the delayed-acceptance ratio, the spread
  does not. Once it exceeds ``surr_tol`` the sampler falls back to plain
  exact steps for the rest of the run. The switch is global and one-way:
  choosing the kernel per walker from its current state alone would break
  detailed balance.
* ``autocorr_time`` / ``ess`` — integrated autocorrelation times (Sokal's
  window, c = 5, on the walker-averaged autocorrelation) and effective
  sample sizes.

Likelihoods are exp(-chi^2 / 2) with the autoscale factor fitted
(re_scale, orchestration.rs:361-388), flat priors on the box.
"""
import math
import time

import numpy as np

from . import Batch, Model
from . import workloads as W
from .surrogate import Mlp

PARAM_NAMES = ("q", "iangle", "r1", "t1", "t2")
# The batched-chi-squared box: q, iangle (deg), r1 (a), t1, t2 (K).
BOX = np.array([[0.1, 1.0], [70.0, 90.0], [0.005, 0.03], [8000.0, 30000.0], [2500.0, 5000.0]])


def to_unit(p, box=BOX):
    return 2.0 * (np.asarray(p, dtype=np.float64) - box[:, 0]) / (box[:, 1] - box[:, 0]) - 1.0


def from_unit(u, box=BOX):
    return box[:, 0] + (np.asarray(u, dtype=np.float64) + 1.0) * 0.5 * (box[:, 1] - box[:, 0])


def in_box(p, box=BOX):
    p = np.atleast_2d(p)
    return np.all((p >= box[:, 0]) & (p <= box[:, 1]), axis=1)


def sample_box(n, seed, box=BOX):
    """n points uniform in the box (splitmix64 stream, portable)."""
    g = W._splitmix64(seed)
    u = np.fromiter((next(g) for _ in range(n * box.shape[0])), dtype=np.float64, count=n * box.shape[0])
    return box[:, 0] + u.reshape(n, box.shape[0]) * (box[:, 1] - box[:, 0])


def surrogate_base_text(nlat=(20, 20, 40, 40)):
    """The one-binary geometry the surrogate and the sampler share: the
    white dwarf + Roche-lobe-filling donor of the single-model
    configuration (q 0.5, iangle 82, r1 0.015, t1 15000 K, t2 3500 K, linear
    limb darkening 0.4 / 0.5), ephemeris t0 = 0, period = 1 so that time is
    phase, with grids of nlat = (nlat1f, nlat1c, nlat2f, nlat2c)."""
    return W.model_text({"q": 0.5, "iangle": 82.0, "r1": 0.015, "t1": 15000.0, "t2": 3500.0,
                         "ldc1_1": 0.4, "ldc2_1": 0.5, "t0": 0.0, "period": 1.0, "pdot": 0.0, "deltat": 0.0,
                         "nlat1f": nlat[0], "nlat1c": nlat[1], "nlat2f": nlat[2], "nlat2c": nlat[3]})


def bin_phases(nbins):
    return (np.arange(nbins) + 0.5) / nbins


def eclipse_phases(n=256, n_dense=192, half_width=0.06):
    """Output phases dense around the white-dwarf eclipse at phase 0:
    n_dense evenly in [-half_width, half_width] (mod 1) and the rest evenly
    over the remaining orbit, sorted in [0, 1)."""
    d = -half_width + (np.arange(n_dense) + 0.5) * (2 * half_width / n_dense)
    m = n - n_dense
    o = half_width + (np.arange(m) + 0.5) * ((1 - 2 * half_width) / m)
    return np.sort(np.mod(np.concatenate([d, o]), 1.0))


def make_training_set(base_text, params, nbins=256, chunk=20000, log=None, phases=None):
    """log(flux / wdwarf) at the output phases (default: nbins bin centres)
    for every parameter row, from the exact GPU method. Rows whose setup
    fails come back as NaN."""
    model = Model.from_text(base_text)
    ph = bin_phases(nbins) if phases is None else np.asarray(phases, dtype=np.float64)
    nbins = len(ph)
    out = np.full((len(params), nbins), np.nan)
    for c0 in range(0, len(params), chunk):
        rows = params[c0:c0 + chunk]
        b = Batch(model, ph, 0.0, 1, names=PARAM_NAMES, values=rows)
        r = b.run(scale=False)
        with np.errstate(invalid="ignore", divide="ignore"):
            out[c0:c0 + len(rows)] = np.log(r["flux"] / r["wdwarf"][:, None])
        del b
        if log:
            log(f"training curves {min(c0 + chunk, len(params))}/{len(params)}")
    return out


def local_box(center, width, nsig=10.0, box=BOX):
    """center +- nsig * width, clipped to the global box."""
    lo = np.maximum(box[:, 0], center - nsig * width)
    hi = np.minimum(box[:, 1], center + nsig * width)
    return np.stack([lo, hi], axis=1)


def quad_features(u):
    """1, u_i, u_i u_j (i <= j) for unit-box coordinates u [n, d]."""
    u = np.atleast_2d(u)
    n, d = u.shape
    cols = [np.ones(n)] + [u[:, i] for i in range(d)]
    cols += [u[:, i] * u[:, j] for i in range(d) for j in range(i, d)]
    return np.stack(cols, axis=1)


class Surrogate:
    """Parameters -> log(flux / wdwarf) at nbins phase bins, trained over
    ``box``; its log-likelihood is -inf outside the box.

    The curve is a quadratic polynomial in the box coordinates, fitted by
    FP64 least squares on the host, plus the MLP's correction, which is
    trained on the normalised residual. The network thus carries only what
    the polynomial misses, so its BF16 precision applies to a much smaller
    quantity."""

    def __init__(self, nbins=256, hidden=(512, 512, 512, 512), seed=1, max_batch=4096, box=BOX, phases=None):
        self.box = np.array(box, dtype=np.float64)
        self.phases = bin_phases(nbins) if phases is None else np.sort(np.asarray(phases, dtype=np.float64))
        nbins = len(self.phases)
        self.nbins = nbins
        self.net = Mlp((len(PARAM_NAMES),) + tuple(hidden) + (nbins,), seed=seed, max_batch=max_batch)
        self.coef = np.zeros((1 + 5 + 15, self.nbins))
        self.y_std = 1.0

    def fit(self, params, logcurves, epochs=60, batch=4096, lr=1e-3, lr_min=1e-5, seed=0, log=None):
        ok = np.all(np.isfinite(logcurves), axis=1)
        u = to_unit(params[ok], self.box)
        y = logcurves[ok]
        F = quad_features(u)
        self.coef = np.linalg.lstsq(F, y, rcond=None)[0]
        r = y - F @ self.coef
        self.y_std = float(r.std()) or 1.0
        self.net.set_data(u.astype(np.float32), (r / self.y_std).astype(np.float32))
        rng = np.random.default_rng(seed)
        hist = []
        t0 = time.perf_counter()
        for e in range(epochs):
            lr_e = lr_min + 0.5 * (lr - lr_min) * (1.0 + math.cos(math.pi * e / max(1, epochs - 1)))
            hist.append(self.net.train_epoch(rng.permutation(len(u)), batch, lr_e))
            if log and (e % 10 == 0 or e == epochs - 1):
                log(f"epoch {e}: loss {hist[-1]:.3e}")
        return {"loss": hist, "seconds": time.perf_counter() - t0, "n_train": int(ok.sum())}

    def predict(self, params):
        """log(flux / wdwarf) [n, nbins]."""
        u = to_unit(params, self.box)
        y = self.net.predict(u.astype(np.float32)).astype(np.float64)
        return quad_features(u) @ self.coef + y * self.y_std

    def curve(self, params, phases):
        """flux / wdwarf at arbitrary phases (periodic linear interpolation
        in log flux between the output phases) [n, len(phases)]."""
        lg = self.predict(params)
        n = self.nbins
        ext = np.concatenate([[self.phases[-1] - 1.0], self.phases, [self.phases[0] + 1.0]])
        x = np.mod(np.asarray(phases, dtype=np.float64), 1.0)
        j = np.clip(np.searchsorted(ext, x, side="right") - 1, 0, n)
        f = (x - ext[j]) / (ext[j + 1] - ext[j])
        i0 = (j - 1) % n
        i1 = j % n
        return np.exp(lg[:, i0] * (1.0 - f) + lg[:, i1] * f)


def autoscaled_chisq(model_flux, y, ferr, w):
    """re_scale (orchestration.rs:361-388) on rows of model curves."""
    wt = w / (ferr * ferr)
    sdy = (model_flux * (wt * y)).sum(1)
    syy = (model_flux * model_flux * wt).sum(1)
    s = np.where(syy > 0, sdy / np.where(syy > 0, syy, 1.0), 1.0)
    r = (y - s[:, None] * model_flux) / ferr
    return (w * r * r).sum(1)


class Target:
    """A synthetic observed curve and its exact / surrogate log-likelihoods."""

    def __init__(self, base_text, truth, times, noise, seed=42, expose=0.0, ndiv=1):
        self.base_text = base_text
        self.model = Model.from_text(base_text)
        self.times = np.asarray(times, dtype=np.float64)
        self.expose, self.ndiv = expose, ndiv
        b = Batch(self.model, self.times, expose, ndiv, names=PARAM_NAMES, values=np.atleast_2d(truth))
        f = b.run(scale=False)["flux"][0]
        rng = np.random.default_rng(seed)
        self.ferr = noise * np.abs(f)
        self.y = f + self.ferr * rng.standard_normal(len(f))
        self.w = np.ones_like(f)
        self.n_exact = 0
        self.t_exact = 0.0

    def logp_exact(self, params):
        params = np.atleast_2d(params)
        out = np.full(len(params), -np.inf)
        if len(params) == 0:
            return out
        t0 = time.perf_counter()
        b = Batch(self.model, self.times, self.expose, self.ndiv, names=PARAM_NAMES, values=params,
                  flux=self.y, ferr=self.ferr, weight=self.w)
        chi = b.run(scale=True, want_flux=False)["chisq"]
        del b
        self.t_exact += time.perf_counter() - t0
        self.n_exact += len(params)
        ok = np.isfinite(chi)
        out[ok] = -0.5 * chi[ok]
        return out

    def logp_surrogate(self, surr, params):
        params = np.atleast_2d(params)
        out = np.full(len(params), -np.inf)
        ins = in_box(params, surr.box)
        if ins.any():
            out[ins] = -0.5 * autoscaled_chisq(surr.curve(params[ins], self.times), self.y, self.ferr, self.w)
        return out


def ensemble_mcmc(target, p0, nsteps, seed=0, surrogate=None, a=2.0, surr_tol=25.0, inner_steps=1,
                  box=BOX):
    """Stretch-move ensemble; delayed acceptance / surrogate transition when
    ``surrogate`` is given. Returns dict with chain [nsteps, nwalkers,
    ndim], acceptance counts, exact-evaluation count and wall time."""
    rng = np.random.default_rng(seed)
    x = np.array(p0, dtype=np.float64)
    nw, nd = x.shape
    if nw % 2:
        raise ValueError("an even number of walkers is needed")
    half = nw // 2
    n_exact0 = target.n_exact
    t0 = time.perf_counter()
    lp = target.logp_exact(x)
    ls = target.logp_surrogate(surrogate, x) if surrogate is not None else None
    chain = np.zeros((nsteps, nw, nd))
    acc = 0
    stage1 = 0
    fallback_at = None
    use_surr = surrogate is not None
    for step in range(nsteps):
        for h in (0, 1):
            k = np.arange(h * half, (h + 1) * half)
            other = np.arange((1 - h) * half, (2 - h) * half)
            j = other[rng.integers(0, half, size=half)]
            z = ((a - 1.0) * rng.random(half) + 1.0) ** 2 / a
            y = x[j] + z[:, None] * (x[k] - x[j])
            logq = (nd - 1) * np.log(z)
            u1 = np.log(rng.random(half))
            u2 = np.log(rng.random(half))
            inside = in_box(y, box)
            lpy = np.full(half, -np.inf)
            if not use_surr:
                cand = inside
                lpy[cand] = target.logp_exact(y[cand])
                ok = cand & (u1 < logq + lpy - lp[k])
            else:
                # Surrogate-only stretch moves of the walkers in k; the
                # complementary half stays fixed throughout.
                xs = x[k].copy()
                lss = ls[k].copy()
                for it in range(inner_steps):
                    if it:
                        j = other[rng.integers(0, half, size=half)]
                        z = ((a - 1.0) * rng.random(half) + 1.0) ** 2 / a
                        y = x[j] + z[:, None] * (xs - x[j])
                        logq = (nd - 1) * np.log(z)
                        u1 = np.log(rng.random(half))
                        inside = in_box(y, box)
                    lsy = np.full(half, -np.inf)
                    lsy[inside] = target.logp_surrogate(surrogate, y[inside])
                    step_ok = inside & (u1 < logq + lsy - lss)
                    xs[step_ok] = y[step_ok]
                    lss[step_ok] = lsy[step_ok]
                moved = np.any(xs != x[k], axis=1)
                stage1 += int(moved.sum())
                lpy[moved] = target.logp_exact(xs[moved])
                ok = moved & (u2 < (lpy - lp[k]) - (lss - ls[k]))
                y = xs
                ls[k[ok]] = lss[ok]
            x[k[ok]] = y[ok]
            lp[k[ok]] = lpy[ok]
            acc += int(ok.sum())
        chain[step] = x
        if use_surr and np.all(np.isfinite(lp)) and np.std(ls - lp) > surr_tol:
            use_surr = False
            fallback_at = step
    return {"chain": chain, "accepted": acc, "proposals": nsteps * nw, "stage1_passed": stage1,
            "fallback_at": fallback_at, "n_exact": target.n_exact - n_exact0,
            "seconds": time.perf_counter() - t0}


def _autocorr_1d(x):
    n = len(x)
    f = np.fft.rfft(x - x.mean(), n=2 * n)
    acf = np.fft.irfft(f * np.conj(f))[:n]
    return acf / acf[0] if acf[0] > 0 else np.ones(n)


def autocorr_time(chain, c=5.0):
    """Integrated autocorrelation time per parameter of a [nsteps, nwalkers,
    ndim] chain: walker-averaged ACF, Sokal's automatic window (M >= c tau)."""
    ns, nw, nd = chain.shape
    tau = np.zeros(nd)
    for d in range(nd):
        acf = np.mean([_autocorr_1d(chain[:, w, d]) for w in range(nw)], axis=0)
        taus = 2.0 * np.cumsum(acf) - 1.0
        m = np.arange(len(taus)) < c * taus
        win = int(np.argmin(m)) if not np.all(m) else len(taus) - 1
        tau[d] = taus[win]
    return tau


def ess(chain, burn=0.25):
    """Effective sample size (min over parameters) after dropping `burn`."""
    c = chain[int(burn * len(chain)):]
    tau = autocorr_time(c)
    return c.shape[0] * c.shape[1] / np.max(tau), tau


def summarize(chain, burn=0.25):
    """Per-parameter 16 / 50 / 84 percentiles after burn-in."""
    c = chain[int(burn * len(chain)):].reshape(-1, chain.shape[2])
    return np.percentile(c, [16, 50, 84], axis=0)
