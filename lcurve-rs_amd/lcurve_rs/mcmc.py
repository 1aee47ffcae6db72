"""Surrogate-assisted MCMC over light-curve models.

* ``make_training_set`` — curves from the exact GPU method over a parameter
  region: log(flux / wdwarf) at phase-bin centres, or at a target's own
  times with its exposure smearing.
* ``Frame`` — the affine coordinates a surrogate works in: the unit box of a
  parameter box, or a ball in whitened coordinates around a posterior (mean
  and covariance of pilot samples), which follows the parameter
  correlations instead of wasting the fit on the corners of a box.
* ``Surrogate`` — an FP64 polynomial baseline in the frame's coordinates
  plus the GPU MLP (lcurve_rs.surrogate.Mlp) trained on its residual, with
  the phase interpolation that turns bins into a curve at arbitrary phases.
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
import itertools
import math
import time

import numpy as np

from . import Batch, Model
from . import workloads as W
from .surrogate import Mlp, SurrogateLikelihood

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


def make_training_set(base_text, params, nbins=256, chunk=20000, log=None, phases=None, times=None, expose=0.0,
                      ndiv=1):
    """log(flux / wdwarf) for every parameter row, from the exact GPU
    method: at the output phases (default: nbins bin centres), or, with
    ``times``, at those times with exposure ``expose`` smeared over ``ndiv``
    sub-times. Rows whose setup fails come back as NaN."""
    model = Model.from_text(base_text)
    if times is not None:
        ph = np.asarray(times, dtype=np.float64)
    else:
        ph = bin_phases(nbins) if phases is None else np.asarray(phases, dtype=np.float64)
        expose, ndiv = 0.0, 1
    nbins = len(ph)
    out = np.full((len(params), nbins), np.nan)
    for c0 in range(0, len(params), chunk):
        rows = params[c0:c0 + chunk]
        b = Batch(model, ph, expose, ndiv, names=PARAM_NAMES, values=rows)
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


def poly_features(u, degree):
    """Every monomial of the coordinates u [n, d] up to ``degree``: 1, u_i,
    u_i u_j (i <= j), u_i u_j u_k (i <= j <= k), ..."""
    u = np.atleast_2d(np.asarray(u, dtype=np.float64))
    n, d = u.shape
    cols = [np.ones(n)]
    for k in range(1, degree + 1):
        cols += [np.prod(u[:, list(c)], axis=1) for c in itertools.combinations_with_replacement(range(d), k)]
    return np.stack(cols, axis=1)


def quad_features(u):
    """1, u_i, u_i u_j (i <= j) for unit-box coordinates u [n, d]."""
    return poly_features(u, 2)


class Frame:
    """Affine coordinates u = (p - center) @ M with a domain: the unit cube
    |u_i| <= 1 (``Frame.from_box``) or the unit ball |u| <= 1
    (``Frame.local``)."""

    def __init__(self, center, M, kind, radius=1.0):
        if kind not in ("box", "ball"):
            raise ValueError("kind must be 'box' or 'ball'")
        self.center = np.asarray(center, dtype=np.float64)
        self.M = np.asarray(M, dtype=np.float64)
        self.Minv = np.linalg.inv(self.M)
        self.kind = kind
        self.radius = float(radius)  # ball: standard deviations per unit of u

    @classmethod
    def from_box(cls, box):
        box = np.asarray(box, dtype=np.float64)
        return cls(0.5 * (box[:, 0] + box[:, 1]), np.diag(2.0 / (box[:, 1] - box[:, 0])), "box")

    @classmethod
    def local(cls, samples, radius=6.0):
        """The ball of ``radius`` posterior standard deviations (Mahalanobis
        distance under the samples' covariance) around the samples' mean."""
        s = np.asarray(samples, dtype=np.float64)
        L = np.linalg.cholesky(np.cov(s, rowvar=False))
        return cls(s.mean(0), np.linalg.inv(L).T / radius, "ball", radius)

    def to_u(self, p):
        return (np.atleast_2d(np.asarray(p, dtype=np.float64)) - self.center) @ self.M

    def from_u(self, u):
        return self.center + np.atleast_2d(u) @ self.Minv

    def contains(self, p):
        u = self.to_u(p)
        if self.kind == "box":
            return np.all(np.abs(u) <= 1.0, axis=1)
        return np.einsum("ij,ij->i", u, u) <= 1.0

    def sample(self, n, seed, core=1.0 / 3.0):
        """n points of the domain. Box: uniform (splitmix64 stream). Ball:
        half from a normal of standard deviation ``core`` (in units of the
        radius) cut to the ball, where the posterior mass lies, and the
        rest uniform in the ball."""
        d = len(self.center)
        if self.kind == "box":
            lo, hi = self.from_u(-np.ones(d))[0], self.from_u(np.ones(d))[0]
            return sample_box(n, seed, np.stack([lo, hi], axis=1))
        rng = np.random.default_rng(seed)
        g = core * rng.standard_normal((2 * (n // 2) + 8, d))
        g = g[np.einsum("ij,ij->i", g, g) <= 1.0][: n // 2]
        m = n - len(g)
        v = rng.standard_normal((m, d))
        v *= (rng.random((m, 1)) ** (1.0 / d)) / np.linalg.norm(v, axis=1)[:, None]
        return self.from_u(np.concatenate([g, v]))


class Surrogate:
    """Parameters -> log(flux / wdwarf) at its output points, trained over
    a ``Frame``; its log-likelihood is -inf outside the frame's domain.

    The outputs are phase bins (``phases``, interpolated to any phase by
    ``curve``) or a target's own times (``times``: ``curve`` then returns
    exactly those points, exposure smearing included when the training
    curves had it). The curve is a polynomial of ``degree`` in the frame
    coordinates, fitted by FP64 least squares on the host, plus the MLP's
    correction, which is trained on the normalised residual. The network
    thus carries only what the polynomial misses, so its BF16 precision
    applies to a much smaller quantity. ``box`` gives a box frame when no
    ``frame`` is passed."""

    def __init__(self, nbins=256, hidden=(512, 512, 512, 512), seed=1, max_batch=4096, box=BOX, phases=None,
                 frame=None, times=None, degree=2, tail=None):
        self.frame = Frame.from_box(box) if frame is None else frame
        if tail and self.frame.kind != "ball":
            raise ValueError("a Gaussian tail needs a ball frame")
        # Outside a ball frame's domain the log-likelihood is, with ``tail``
        # (the default for ball frames), the Gaussian of the frame's
        # covariance: the value at the centre minus half the squared
        # Mahalanobis distance. A positive surrogate density everywhere keeps
        # delayed acceptance exact on the whole prior. Without a tail (box
        # frames, or tail=False) it is -inf there, and the sampler then
        # targets the posterior restricted to the surrogate's domain.
        self.tail = (self.frame.kind == "ball") if tail is None else bool(tail)
        self.degree = int(degree)
        self.times = None if times is None else np.array(times, dtype=np.float64)
        if self.times is not None:
            self.phases = None
            nbins = len(self.times)
        else:
            self.phases = bin_phases(nbins) if phases is None else np.sort(np.asarray(phases, dtype=np.float64))
            nbins = len(self.phases)
        self.nbins = nbins
        # The MLP's layer widths are multiples of 64; extra outputs train to 0.
        self.nout = -(-nbins // 64) * 64
        nd = len(self.frame.center)
        self.net = Mlp((nd,) + tuple(hidden) + (self.nout,), seed=seed, max_batch=max_batch)
        self.coef = np.zeros((poly_features(np.zeros((1, nd)), self.degree).shape[1], self.nbins))
        self.y_std = 1.0
        self.version = 0  # bumped by every fit

    @property
    def box(self):
        """The parameter box of a box frame."""
        if self.frame.kind != "box":
            raise AttributeError("a ball frame has no box")
        d = len(self.frame.center)
        return np.stack([self.frame.from_u(-np.ones(d))[0], self.frame.from_u(np.ones(d))[0]], axis=1)

    def contains(self, params):
        return self.frame.contains(params)

    def fit(self, params, logcurves, epochs=60, batch=4096, lr=1e-3, lr_min=1e-5, seed=0, log=None):
        ok = np.all(np.isfinite(logcurves), axis=1)
        u = self.frame.to_u(params[ok])
        y = logcurves[ok]
        F = poly_features(u, self.degree)
        self.version += 1
        self.coef = np.linalg.lstsq(F, y, rcond=None)[0]
        r = y - F @ self.coef
        self.y_std = float(r.std()) or 1.0
        t = np.zeros((len(u), self.nout), dtype=np.float32)
        t[:, :self.nbins] = r / self.y_std
        self.net.set_data(u.astype(np.float32), t)
        rng = np.random.default_rng(seed)
        hist = []
        t0 = time.perf_counter()
        for e in range(epochs):
            lr_e = lr_min + 0.5 * (lr - lr_min) * (1.0 + math.cos(math.pi * e / max(1, epochs - 1)))
            hist.append(self.net.train_epoch(rng.permutation(len(u)), batch, lr_e))
            if log and (e % 10 == 0 or e == epochs - 1):
                log(f"epoch {e}: loss {hist[-1]:.3e}")
        return {"loss": hist, "seconds": time.perf_counter() - t0, "n_train": int(ok.sum())}

    def likelihood(self, y, ferr, w):
        """The autoscaled log-likelihood of observed fluxes y (at this
        surrogate's times) on the GPU: a callable params [n, d] -> [n]. It
        ignores the domain; callers mask rows outside it."""
        if self.times is None:
            raise ValueError("the GPU likelihood needs a surrogate of the data times")
        return SurrogateLikelihood(self.net, self.frame.center, self.frame.M, self.degree, self.coef, self.y_std,
                                   y, ferr, w)

    def baseline(self, params):
        """The polynomial alone [n, nbins]."""
        return poly_features(self.frame.to_u(params), self.degree) @ self.coef

    def predict(self, params):
        """log(flux / wdwarf) [n, nbins]."""
        u = self.frame.to_u(params)
        y = self.net.predict(u.astype(np.float32))[:, :self.nbins].astype(np.float64)
        return poly_features(u, self.degree) @ self.coef + y * self.y_std

    def curve(self, params, phases):
        """flux / wdwarf [n, len(phases)]: at the output times themselves
        (a time-output surrogate only serves its own times), or by periodic
        linear interpolation in log flux between the output phases."""
        lg = self.predict(params)
        if self.times is not None:
            if not np.array_equal(np.asarray(phases, dtype=np.float64), self.times):
                raise ValueError("this surrogate predicts its training times only")
            return np.exp(lg)
        n = self.nbins
        ext = np.concatenate([[self.phases[-1] - 1.0], self.phases, [self.phases[0] + 1.0]])
        x = np.mod(np.asarray(phases, dtype=np.float64), 1.0)
        j = np.clip(np.searchsorted(ext, x, side="right") - 1, 0, n)
        f = (x - ext[j]) / (ext[j + 1] - ext[j])
        i0 = (j - 1) % n
        i1 = j % n
        return np.exp(lg[:, i0] * (1.0 - f) + lg[:, i1] * f)


def autoscaled_chisq(model_flux, y, ferr, w):
    """re_scale (orchestration.rs:361-388) on rows of model curves: points
    with weight <= 0 take no part, and a row with no such point or with
    sum w m^2 / ferr^2 = 0 gets chi-squared 0 (the reference's (1, 0, 0))."""
    model_flux = np.atleast_2d(model_flux)
    ok = w > 0
    yo, eo, wo = y[ok], ferr[ok], w[ok]
    mo = model_flux[:, ok]
    wt = wo / (eo * eo)
    sdy = (wt * yo * mo).sum(1)
    syy = (wt * mo * mo).sum(1)
    good = (wo.sum() > 0) & (syy > 0)
    s = np.where(good, sdy / np.where(good, syy, 1.0), 1.0)
    r = (yo - mo * s[:, None]) / eo
    return np.where(good, (wo * (r * r)).sum(1), 0.0)


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
        self._lik = None  # (surrogate, its version, GPU likelihood)

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
        """Surrogate log-likelihood: -inf outside its domain, or its
        Gaussian tail there (Surrogate ``tail``). A surrogate of this
        target's own times is evaluated on the GPU in one call
        (Surrogate.likelihood); others through their interpolated curve."""
        return self._logp_surrogate(surr, params, gpu=True)

    def logp_surrogate_host(self, surr, params):
        """The same on the host from the surrogate's curve (reference for
        the GPU likelihood)."""
        return self._logp_surrogate(surr, params, gpu=False)

    def _logp_surrogate(self, surr, params, gpu):
        params = np.atleast_2d(params)
        out = np.full(len(params), -np.inf)
        ins = surr.contains(params)
        own_times = getattr(surr, "times", None) is not None and np.array_equal(surr.times, self.times)

        def inner(p):
            if gpu and own_times:
                if self._lik is None or self._lik[0] is not surr or self._lik[1] != surr.version:
                    c = surr.frame.center[None]
                    lik = surr.likelihood(self.y, self.ferr, self.w)
                    self._lik = (surr, surr.version, lik, float(lik(c)[0]))
                return self._lik[2](p)
            return -0.5 * autoscaled_chisq(surr.curve(p, self.times), self.y, self.ferr, self.w)

        if ins.any():
            out[ins] = inner(params[ins])
        if getattr(surr, "tail", False) and not ins.all():
            if gpu and own_times:
                inner(params[:0])  # make sure the centre value is cached
                l0 = self._lik[3]
            else:
                l0 = float(inner(surr.frame.center[None])[0])
            u = surr.frame.to_u(params[~ins])
            out[~ins] = l0 - 0.5 * surr.frame.radius ** 2 * np.einsum("ij,ij->i", u, u)
        return out


def ensemble_mcmc(target, p0, nsteps, seed=0, surrogate=None, a=2.0, surr_tol=25.0, inner_steps=1,
                  box=BOX, log=None, log_every=50):
    """Stretch-move ensemble; delayed acceptance / surrogate transition when
    ``surrogate`` is given. Returns dict with chain [nsteps, nwalkers,
    ndim], acceptance counts, exact-evaluation count and wall time.
    ``log`` (a callable) receives a progress line every ``log_every``
    steps. With a surrogate, every initial walker must have a finite
    surrogate log-likelihood (inside its domain, or any point for a
    surrogate with a tail); a surrogate without a tail confines the chain to
    its domain, so the chain then samples the posterior restricted to it.
    Delayed acceptance falls back to exact steps for the rest of the run when
    the spread of surrogate - exact log-likelihood over the walkers exceeds
    ``surr_tol`` or is not finite."""
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
    if ls is not None and not np.all(np.isfinite(ls)):
        raise ValueError(f"{int(np.sum(~np.isfinite(ls)))} initial walkers have no finite surrogate "
                         "log-likelihood (outside the surrogate's domain, which has no tail)")
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
        if use_surr and np.all(np.isfinite(lp)) and not (np.std(ls - lp) <= surr_tol):
            use_surr = False
            fallback_at = step
        if log and (step + 1) % log_every == 0:
            log(f"step {step + 1}/{nsteps}: acceptance {acc / ((step + 1) * nw):.3f}, "
                f"exact evaluations {target.n_exact - n_exact0}, {time.perf_counter() - t0:.1f} s")
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
