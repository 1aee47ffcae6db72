"""Pins of the oracle's functions that the reference's golden curve does not
reach, each against mathematics that does not restate the oracle's code.

Each test names the reference function it pins and the mistakes it would
catch. The checks come from three kinds of source:
  * closed forms: a flat disc of uniform or power-law brightness
    (flux.rs:83-91 with brightness.rs:204-212), a cylindrical rim
    (flux.rs:92-100), the bright-spot strip as a geometric series
    (flux.rs:101-106, grid.rs:484-553), the first-order Doppler beaming
    derivative of a sphere (flux.rs:40-47, 67-74, 138-146, 181-188), and
    the surface gravity and radius of spheres (flux.rs:282-375);
  * textbook formulas: the point-lens magnification of the major image
    (flux.rs:390-406), the Roche potential and its gradient for surface
    positions, normals and gravities (surface.rs:54-121, grid.rs:166-181),
    Eggleton's Roche-lobe radius, the Gaussian star-spot and tapered
    equatorial-band temperatures (brightness.rs:13-147), and the Roemer
    light-travel delay of the phases (orchestration.rs:231-237);
  * independent numerics: the exposure average as a converging integral
    (flux.rs:378-387, 110), the gas stream by scipy's ODE solver from the L1
    point's unstable direction (stream.rs:8-120), and the lobe volume by
    root finding along rays.
"""
import math

import numpy as np
import pytest
from scipy.integrate import quad, solve_ivp
from scipy.optimize import brentq

import pyoracle
from lcurve_rs import workloads as W
from test_oracle import planck_nu, sphere_flux, sphere_model

G_SI, C_SI, DAY = 6.673e-11, 2.99792458e8, 86400.0   # crates/subs/src/constants.rs
H_SI, K_SI = 6.6262e-34, 1.3806e-23
C_KMS = C_SI / 1e3


def earth_vec(iangle_deg, phase):
    i = math.radians(iangle_deg)
    p = 2.0 * math.pi * np.asarray(phase, dtype=float)
    return np.stack([math.sin(i) * np.cos(p), -math.sin(i) * np.sin(p), np.full(p.shape, math.cos(i))], -1)


def kepler(vscale_kms, tperiod_days):
    """Orbital separation a (m) and G(M1+M2) (SI) from the relative orbital
    speed and period: v = 2 pi a / P, G M = (2 pi / P)^2 a^3."""
    p = tperiod_days * DAY
    a = 1000.0 * vscale_kms * p / (2.0 * math.pi)
    return a, (2.0 * math.pi / p) ** 2 * a ** 3


def brightness_temperature(b, wave_nm=550.0):
    """Inverse of Planck's B_nu at wave_nm."""
    lam = wave_nm * 1e-9
    return H_SI * C_SI / (lam * K_SI) / np.log1p(2.0 * H_SI * C_SI / lam ** 3 / b)


def _roche(q):
    """Synchronous Roche potential (units G(M1+M2)/a, star 1 at the origin,
    star 2 at (1,0,0)), its gradient and the inner Lagrange point."""
    mu = q / (1.0 + q)

    def pot(p):
        x, y, z = p[..., 0], p[..., 1], p[..., 2]
        return (-(1 - mu) / np.sqrt(x * x + y * y + z * z) - mu / np.sqrt((x - 1) ** 2 + y * y + z * z)
                - 0.5 * ((x - mu) ** 2 + y * y))

    def grad(p):
        x, y, z = p[..., 0], p[..., 1], p[..., 2]
        r1 = np.sqrt(x * x + y * y + z * z)
        r2 = np.sqrt((x - 1) ** 2 + y * y + z * z)
        gx = (1 - mu) * x / r1 ** 3 + mu * (x - 1) / r2 ** 3 - (x - mu)
        gy = (1 - mu) * y / r1 ** 3 + mu * y / r2 ** 3 - y
        gz = (1 - mu) * z / r1 ** 3 + mu * z / r2 ** 3
        return np.stack([gx, gy, gz], -1)

    xl1 = brentq(lambda x: grad(np.array([x, 0.0, 0.0]))[0], 1e-6, 1 - 1e-6, xtol=1e-15)
    return pot, grad, xl1, mu


NO_ECL = {"eclipse1": 0, "eclipse2": 0}


# ------------------------------------------------------ disc (flux.rs:83-91)
def disc_model(texp=0.0, nrad=60, iangle=60.0, lin=0.3, quad_=0.2, height=0.0, beta=1.5):
    ov = {"add_disc": 1, "opaque": 0, "height_disc": height, "beta_disc": beta, "rdisc1": 0.1,
          "rdisc2": 0.3, "nrad": nrad, "temp_disc": 6000.0, "texp_disc": texp, "lin_limb_disc": lin,
          "quad_limb_disc": quad_, "temp_edge": 5000.0, "absorb_edge": 0.0, "iangle": iangle,
          "nlat1f": 12, "nlat1c": 12, "nlat2f": 16, "nlat2c": 16}
    ov.update(NO_ECL)
    return W.model_text(ov)


def disc_ld(mu, lin=0.3, quad_=0.2):
    return 1.0 - (1.0 - mu) * (lin + quad_ * (1.0 - mu))


def light_minus(txt, off, times, **kw):
    """comp_light of a model minus that of the same model with `off` applied:
    the light of the component that `off` removes."""
    a, _ = pyoracle.light_curve(txt, times, nthreads=4, **kw)
    b, _ = pyoracle.light_curve(W.model_text(off, txt), times, nthreads=4, **kw)
    return a - b


@pytest.mark.parametrize("iangle", [30.0, 60.0, 85.0])
def test_flat_uniform_disc_is_cos_i_ld_b_times_area(iangle):
    """A flat disc (height 0: every normal is +z, mu = cos i) of uniform
    brightness (texp 0): flux = cos i LD(cos i) B(T) pi (r2^2 - r1^2),
    whatever the phase, in comp_light and comp_disc. Catches a wrong sign or
    a dropped term of the quadratic disc law, a missing mu factor, and ring
    areas that do not tile the annulus."""
    txt = disc_model(iangle=iangle)
    t = np.linspace(0.05, 0.95, 7)
    mu = math.cos(math.radians(iangle))
    expect = mu * disc_ld(mu) * planck_nu(550.0, 6000.0) * math.pi * (0.3 ** 2 - 0.1 ** 2)
    got = light_minus(txt, {"add_disc": 0}, t)
    # element fluxes and areas are FP32: the sum carries ~1e-7 relative error
    assert np.all(np.abs(got / expect - 1.0) < 3e-7), got / expect - 1
    fcomp, _ = pyoracle.components(txt, t, nthreads=4)
    assert np.all(np.abs(fcomp[:, 1] / expect - 1.0) < 3e-7)


def test_power_law_disc_converges_to_the_radial_integral():
    """Brightness B(T) (r / rdisc2)^texp (brightness.rs:204-212): the disc
    flux is cos i LD(cos i) times the integral of 2 pi r B(r) dr, reached at
    second order in the ring width (rings at their mid radii)."""
    texp, mu = -0.75, math.cos(math.radians(60.0))
    b = planck_nu(550.0, 6000.0)
    integral = 2 * math.pi * b * 0.3 ** (-texp) * (0.3 ** (texp + 2) - 0.1 ** (texp + 2)) / (texp + 2)
    expect = mu * disc_ld(mu) * integral
    errs = []
    for nrad in (30, 60):
        fcomp, _ = pyoracle.components(disc_model(texp=texp, nrad=nrad), [0.3, 0.4], nthreads=4)
        errs.append(fcomp[0, 1] / expect - 1.0)
    assert abs(errs[1]) < 2e-5, errs
    assert 3.0 < errs[0] / errs[1] < 5.0, errs


def test_disc_rim_is_a_cylinder():
    """The outer rim (flux.rs:92-100; vertical elements with horizontal
    normals, uniform brightness when absorb_edge = 0) seen at inclination i:
    flux = B (sum of rim areas) / (2 pi) * integral over the visible half of
    sin i cos t LD(sin i cos t) dt, at every phase."""
    iangle = 70.0
    txt = disc_model(height=0.05, beta=0.0, nrad=40, iangle=iangle)
    t = np.array([0.0, 0.13, 0.37, 0.71])
    ob = pyoracle.OracleBatch(txt, [0.0], 0.001, 1)
    area = ob.grid(0, 5)["area"].sum()
    s = math.sin(math.radians(iangle))
    ang = quad(lambda th: s * math.cos(th) * disc_ld(s * math.cos(th)), -math.pi / 2, math.pi / 2,
               epsabs=0, epsrel=1e-13)[0]
    expect = planck_nu(550.0, 5000.0) * area / (2 * math.pi) * ang
    fcomp, _ = pyoracle.components(txt, t, nthreads=4)
    # the rim is sampled at ~250 longitudes, and the visible half ends at a
    # kink of the integrand: a few 1e-5 of discretisation error
    assert np.all(np.abs(fcomp[:, 2] / expect - 1.0) < 2e-4), fcomp[:, 2] / expect - 1
    top = pyoracle.components(disc_model(height=0.05, beta=0.0, nrad=40, iangle=iangle), t, nthreads=4)[0][:, 1]
    got = light_minus(txt, {"add_disc": 0}, t)
    assert np.allclose(got, fcomp[:, 2] + top, rtol=1e-12)


# -------------------------------------- bright spot (flux.rs:101-106)
SPOT = {"add_spot": 1, "nspot": 200, "expon_spot": 0.0, "epow_spot": 1.0, "radius_spot": 0.3,
        "length_spot": 0.02, "height_spot": 0.05, "temp_spot": 15000.0, "angle_spot": 140.0,
        "yaw_spot": 10.0, "tilt_spot": 60.0, "cfrac_spot": 0.2, "nlat1f": 12, "nlat1c": 12,
        "nlat2f": 16, "nlat2c": 16}


def _stream_point(q, radius):
    """Ballistic stream from L1 to the first point at `radius` from star 1:
    rotating-frame equations of motion (Coriolis + centrifugal + both
    point masses), started 1e-5 inside L1 along the unstable eigenvector of
    the linearised motion there, integrated by scipy's DOP853."""
    mu = q / (1.0 + q)

    def acc(p):
        x, y = p
        r1 = math.hypot(x, y)
        r2 = math.hypot(x - 1.0, y)
        return np.array([-(1 - mu) * x / r1 ** 3 - mu * (x - 1) / r2 ** 3 + (x - mu),
                         -(1 - mu) * y / r1 ** 3 - mu * y / r2 ** 3 + y])

    def rhs(_, s):
        a = acc(s[:2])
        return [s[2], s[3], a[0] + 2.0 * s[3], a[1] - 2.0 * s[2]]

    _, _, xl1, _ = _roche(q)
    h = 1e-7
    jac = np.zeros((4, 4))
    jac[0, 2] = jac[1, 3] = 1.0
    jac[2, 3], jac[3, 2] = 2.0, -2.0
    for k in range(2):
        dp = np.zeros(2)
        dp[k] = h
        jac[2:, k] = (acc(np.array([xl1, 0.0]) + dp) - acc(np.array([xl1, 0.0]) - dp)) / (2 * h)
    w, v = np.linalg.eig(jac)
    k = int(np.argmax(w.real))
    vec = v[:, k].real
    vec = vec / (-vec[0]) * 1e-5          # displacement 1e-5 towards star 1
    s0 = np.array([xl1, 0.0, 0.0, 0.0]) + vec
    ev = lambda _, s: math.hypot(s[0], s[1]) - radius
    ev.terminal, ev.direction = True, -1
    sol = solve_ivp(rhs, (0.0, 10.0), s0, method="DOP853", rtol=1e-12, atol=1e-14, events=ev)
    return sol.y_events[0][0][:2]


def test_bright_spot_strip_geometry_and_flux():
    """Bright spot with expon 0, epow 1: element i of the strip sits at
    bspot + L dist_i bvec with brightness B(T) exp(-dist_i), dist_i = 20 i /
    (n - 1), so the strip's total is the geometric series
    B L H h (1 - e^(-n h)) / (1 - e^(-h)), h = 20 / (n - 1); a fraction
    cfrac faces up, the rest along the tilted normal. bspot is where the
    stream from L1 reaches radius_spot. Flux at phase p = total * ((1 -
    cfrac) max(0, e.t) + cfrac max(0, cos i)). Catches the two strips'
    fractions swapped, a wrong tilt or yaw axis, a spot away from the
    stream, and a wrong brightness fall-off."""
    q, iangle = 0.5, 82.0
    txt = W.model_text(dict(SPOT, **NO_ECL))
    ob = pyoracle.OracleBatch(txt, [0.0], 0.001, 1)
    g = ob.grid(0, 6)
    n = SPOT["nspot"]
    posn, dirn, flux = g["posn"], g["dirn"], g["flux"]
    assert len(flux) == 2 * n
    bspot = posn[0]
    assert abs(np.hypot(bspot[0], bspot[1]) - 0.3) < 1e-8 and bspot[2] == 0.0
    assert np.allclose(bspot[:2], _stream_point(q, 0.3), atol=2e-6)
    bvec = (posn[-1] - posn[0]) / np.linalg.norm(posn[-1] - posn[0])
    theta = math.atan2(bspot[1], bspot[0]) + math.pi / 2 + math.radians(140.0)
    assert np.allclose(bvec, [math.cos(theta), math.sin(theta), 0.0], atol=1e-12)
    tilt, yaw = math.radians(60.0), math.radians(10.0)
    tvec = np.array([math.sin(tilt) * math.sin(theta + yaw), -math.sin(tilt) * math.cos(theta + yaw),
                     math.cos(tilt)])
    assert np.allclose(dirn[0::2], tvec, atol=1e-12) and np.allclose(dirn[1::2], [0, 0, 1], atol=0)
    hstep = 20.0 / (n - 1)
    total = (planck_nu(550.0, 15000.0) * 0.02 * 0.05 * hstep * (1 - math.exp(-n * hstep))
             / (1 - math.exp(-hstep)))
    assert abs(flux.sum() / total - 1) < 2e-7
    assert abs(flux[1::2].sum() / (0.2 * total) - 1) < 2e-7
    t = np.linspace(0.0, 1.0, 9)
    e = earth_vec(iangle, t)
    expect = total * (0.8 * np.maximum(e @ tvec, 0.0) + 0.2 * math.cos(math.radians(iangle)))
    got = light_minus(txt, {"add_spot": 0}, t)
    assert np.allclose(got, expect, rtol=3e-7, atol=0)
    fcomp, _ = pyoracle.components(txt, t, nthreads=4)
    assert np.allclose(fcomp[:, 3], expect, rtol=3e-7, atol=0)


# --------------------------- Doppler beaming (flux.rs:40-47, 67-74, 138-146)
def beam_model(b1=0.0, b2=0.0, v=0.0, spin1=1.0, spin2=1.0, nlat=80):
    txt = sphere_model(ld={"ldc1_1": 0.4, "ldc2_1": 0.5}, nlat=nlat, iangle=70.0)
    return W.model_text({"velocity_scale": v, "beam_factor1": b1, "beam_factor2": b2, "spin1": spin1,
                         "spin2": spin2}, txt)


@pytest.mark.parametrize("spin", [1.0, 3.0])
def test_beaming_first_order_is_the_orbital_radial_velocity(spin):
    """To first order in v/c a uniformly bright, linearly limb-darkened
    sphere moving with the orbit changes flux by -b v_r / c: star 1 moves at
    -q/(1+q) v_orb along y, star 2 at +1/(1+q) v_orb. The star's rotation
    adds nothing at first order (its radial velocities cancel over the
    visible hemisphere, and the aberration terms of the limb darkening
    cancel against the radial-velocity term). Central difference in
    velocity_scale against that derivative, for comp_light (with spins) and
    the single-star comp_star1 / comp_star2 (spin-free form). Catches a
    wrong sign of either star's velocity, a missing (1 - b v_r) factor and
    a velocity that is not divided by c."""
    q, ia, dv = 0.5, 70.0, 40.0
    xc = q / (1 + q)
    t = np.array([0.1, 0.25, 0.6, 0.75])
    ey = -math.sin(math.radians(ia)) * np.sin(2 * np.pi * t)
    f1 = sphere_flux(0.05, 15000.0, 0.5 - 0.4 / 6)
    f2 = sphere_flux(0.2, 4000.0, 0.5 - 0.5 / 6)
    cases = [(dict(b1=1.5), -1.5 * ey * xc * f1 / C_KMS, 0),
             (dict(b2=1.2), 1.2 * ey * (1 - xc) * f2 / C_KMS, 4)]
    for kw, deriv, col in cases:
        hi = beam_model(v=dv, spin1=spin, spin2=spin, **kw)
        lo = beam_model(v=-dv, spin1=spin, spin2=spin, **kw)
        fd = (pyoracle.light_curve(hi, t, nthreads=4)[0] - pyoracle.light_curve(lo, t, nthreads=4)[0]) / (2 * dv)
        assert np.allclose(fd, deriv, rtol=2e-3, atol=0), (kw, fd / deriv - 1)
        fdc = (pyoracle.components(hi, t, nthreads=4)[0][:, col]
               - pyoracle.components(lo, t, nthreads=4)[0][:, col]) / (2 * dv)
        assert np.allclose(fdc, deriv, rtol=2e-3, atol=0), (kw, fdc / deriv - 1)


# ------------------------------------- point-lens magnification (flux.rs:390-406)
def major_image_magnification(u):
    """Magnification of the outer (major) image of a point lens at impact
    parameter u Einstein radii: (u^2 + 2) / (2 u sqrt(u^2 + 4)) + 1/2."""
    return (u * u + 2.0) / (2.0 * u * np.sqrt(u * u + 4.0)) + 0.5


def test_lensing_mag_is_the_major_image_magnification():
    """lensing_mag against the textbook: 1 in front of the lens; behind it,
    the major-image magnification at u = p / sqrt(rlens d), exactly within
    10 Einstein radii and to its large-u expansion (1 + u^-4, error
    < 3 u^-6) beyond, where the reference switches to a series form."""
    rng = np.random.default_rng(3)
    near = far = 0
    for it in range(4000):
        rlens = (1e-4, 1e-2)[it % 2]
        e = rng.normal(size=3)
        e /= np.linalg.norm(e)
        s = rng.normal(size=3) * rng.uniform(0.01, 0.4)
        d = -s @ e
        m = pyoracle.lensing_mag(s, e, rlens)
        if d <= 0:
            assert m == 1.0
            continue
        u = np.linalg.norm(s + d * e) / math.sqrt(rlens * d)
        a = major_image_magnification(u)
        if u <= 10.0:
            near += 1
            assert abs(m / a - 1) < 1e-12, (u, m, a)
        else:
            far += 1
            assert abs(m - a) < 3.0 * u ** -6 + 4e-16, (u, m, a)
    assert near > 100 and far > 100


def test_lensed_light_curve_matches_textbook_magnification():
    """Star 2 behind the white dwarf (edge-on, phase near 0.5, eclipses off
    so only the magnification acts): the light gained is the sum over star 2's
    elements behind star 1 of their light times (A(u) - 1), with the
    Einstein radius from the masses: rlens = 4 G M1 / (c^2 a), R_E^2 =
    rlens d (units of a)."""
    q, vs, tp = 0.5, 3000.0, 0.15
    txt = W.model_text({"glens1": 1, "velocity_scale": vs, "tperiod": tp, "ldc2_1": 0.5, **NO_ECL},
                       sphere_model(nlat=90, r1=0.015, r2=0.3, iangle=90.0))
    t = np.array([0.5, 0.49, 0.475])
    got = light_minus(txt, {"glens1": 0}, t)
    a_m, gm = kepler(vs, tp)
    rlens = 4.0 * gm / (1.0 + q) / (C_SI ** 2 * a_m)
    ob = pyoracle.OracleBatch(txt, [0.0], 0.001, 1)
    g = ob.grid(0, 2)
    e = earth_vec(90.0, t)
    for k in range(len(t)):
        mu = g["dirn"] @ e[k]
        d = -(g["posn"] @ e[k])
        vis = (mu > 0) & (d > 0)
        p = np.linalg.norm(g["posn"][vis] + d[vis, None] * e[k][None, :], axis=1)
        u = p / np.sqrt(rlens * d[vis])
        light = mu[vis] * g["flux"][vis] * (1 - 0.5 * (1 - mu[vis]))
        expect = np.sum(light * (major_image_magnification(u) - 1.0))
        bound = np.sum(light[u > 10] * 3.0 * u[u > 10] ** -6.0) + 1e-12 * np.sum(light)
        assert np.sum(u <= 10) > 20
        assert abs(got[k] - expect) <= bound, (k, got[k], expect, bound)


# ------------------------------------- exposure smearing (flux.rs:378-387)
def test_ndiv_matches_trapezoid_of_sub_time_curves(golden_model_text):
    """On the eclipsing golden model, through both eclipses and the fine /
    coarse switch at phase1 = 0.05: ndiv = 5 equals the trapezoid average
    of unsmeared curves at the five times t + expose (k - 2) / 4 (each of
    which picks its own grid and Ginterp factor)."""
    t = np.concatenate([np.linspace(-0.03, 0.06, 31), np.linspace(0.46, 0.54, 9)])
    ex = 0.02
    f5, _ = pyoracle.light_curve(golden_model_text, t, expose=ex, ndiv=5, nthreads=4)
    sub = np.stack([pyoracle.light_curve(golden_model_text, t + ex * (k - 2) / 4, expose=ex, ndiv=1,
                                         nthreads=4)[0] for k in range(5)])
    w = np.array([0.5, 1, 1, 1, 0.5])[:, None]
    assert np.allclose(f5, (w * sub).sum(0) / 4, rtol=1e-13, atol=0)


def test_smearing_converges_to_the_exposure_average():
    """Exposure smearing approximates the mean of the unsmeared curve over
    [t - e/2, t + e/2]: on a smooth phase-varying curve (a Roche-distorted
    donor, ellipsoidal variation, no eclipses) the error falls as
    1 / (ndiv - 1)^2. Catches a window not centred on the time, end weights
    other than 1/2 and a wrong normalisation, none of which converge."""
    txt = W.model_text({**NO_ECL, "nlat1f": 12, "nlat1c": 12, "nlat2f": 40, "nlat2c": 40})
    t = np.array([0.11, 0.2, 0.33, 0.6])
    ex = 0.08
    # the average by Simpson's rule over 257 unsmeared points
    x = np.linspace(-0.5, 0.5, 257)
    dense = np.stack([pyoracle.light_curve(txt, t + ex * xi, nthreads=4)[0] for xi in x])
    wts = np.ones(257)
    wts[1:-1:2], wts[2:-1:2] = 4, 2
    avg = (wts[:, None] * dense).sum(0) / wts.sum()
    errs = []
    for nd in (3, 5, 9, 17):
        f, _ = pyoracle.light_curve(txt, t, expose=ex, ndiv=nd, nthreads=4)
        errs.append(np.max(np.abs(f / avg - 1)))
    assert errs[0] > 1e-6                      # the curve does vary over the window
    for a, b in zip(errs, errs[1:]):
        assert 3.0 < a / b < 5.0, errs
    assert errs[-1] < errs[0] / 40


# ------------------------------------------ Roemer delay (orchestration.rs:231-237)
def test_roemer_delay_shifts_half_phase_by_deltat():
    """With deltat, light from the far side of the orbit is late by deltat:
    phase 0 is reached at t0 (no delay), phase 0.25 at t0 + P/4 + deltat/2
    and phase 0.5 at t0 + P/2 + deltat, to first order in deltat / P. The
    shift is linear in deltat. Catches a wrong sign, a missing factor 1/2
    and sine in place of cosine."""
    t0, per, dt = 0.3, 0.7, 3e-4
    txt = W.model_text({"t0": t0, "period": per, "deltat": dt})
    ph = lambda tt: pyoracle.phases(txt, np.array([tt]))[0]
    for k in (0, 3):
        assert abs(ph(t0 + k * per) - k) < 1e-14
        for frac, delay in ((0.25, dt / 2), (0.5, dt), (0.75, dt / 2)):
            tt = brentq(lambda x: ph(x) - (k + frac), t0 + (k + frac - 0.1) * per,
                        t0 + (k + frac + 0.1) * per, xtol=1e-15)
            assert abs(tt - (t0 + (k + frac) * per + delay)) < 2e-3 * dt, (k, frac, tt)
    txt2 = W.model_text({"t0": t0, "period": per, "deltat": 2 * dt})
    tt = np.linspace(0.0, 5.0, 37)
    base = pyoracle.phases(W.model_text({"t0": t0, "period": per}), tt)
    s1 = pyoracle.phases(txt, tt) - base
    s2 = pyoracle.phases(txt2, tt) - base
    assert np.allclose(s2, 2 * s1, rtol=1e-9, atol=1e-15)


# ------------------------------ gravity and radius (flux.rs:282-375)
def test_sphere_logg_and_radius():
    """Spheres: log g = log10(G M_i / R_i^2) in cgs with M1 = M / (1 + q),
    M2 = q M / (1 + q), G M and a from Kepler's law, and the volume-averaged
    radius equals the sphere's radius."""
    q, vs, tp = 0.5, 400.0, 0.15
    txt = W.model_text({"velocity_scale": vs, "tperiod": tp, "q": q},
                       sphere_model(nlat=40, r1=0.05, r2=0.2))
    _, sc = pyoracle.light_curve(txt, [0.2, 0.3])
    a_m, gm = kepler(vs, tp)
    g1 = 100.0 * gm / (1 + q) / (0.05 * a_m) ** 2
    g2 = 100.0 * gm * q / (1 + q) / (0.2 * a_m) ** 2
    assert abs(sc["logg1"] - math.log10(g1)) < 1e-9
    assert abs(sc["logg2"] - math.log10(g2)) < 1e-9
    assert abs(sc["rv1"] - 0.05) < 1e-12 and abs(sc["rv2"] - 0.2) < 1e-12


def eggleton(q):
    """Eggleton (1983): volume-equivalent Roche-lobe radius of the star of
    mass ratio q = M_star / M_other, good to 1%."""
    q3 = q ** (1.0 / 3.0)
    return 0.49 * q3 * q3 / (0.6 * q3 * q3 + math.log(1 + q3))


def _lobe_volume_radius(q, n=48):
    """Volume-equivalent radius of star 2's Roche lobe: for each direction
    from star 2 the lobe radius by root finding on Phi = Phi(L1), then
    (3 V / 4 pi)^(1/3) by Gauss-Legendre quadrature in cos(colatitude) about
    the x axis and the trapezoid rule in longitude."""
    pot, _, xl1, _ = _roche(q)
    pl1 = pot(np.array([xl1, 0.0, 0.0]))
    xs, ws = np.polynomial.legendre.leggauss(n)
    vol = 0.0
    for ct, wc in zip(xs, ws):
        st = math.sqrt(1 - ct * ct)
        for ph in np.linspace(0, 2 * np.pi, 2 * n, endpoint=False):
            d = np.array([-ct, st * math.cos(ph), st * math.sin(ph)])   # colatitude from -x (L1)
            rmax = 1.0 - xl1
            f = lambda r: pot(np.array([1.0, 0, 0]) + r * d) - pl1
            r = rmax if f(rmax) <= 0 else brentq(f, 1e-3, rmax, xtol=1e-13)
            vol += wc * (np.pi / n) * r ** 3 / 3.0
    return (3.0 * vol / (4.0 * np.pi)) ** (1.0 / 3.0)


@pytest.mark.parametrize("q", [0.2, 0.5, 1.0])
def test_roche_lobe_radius_and_surface(q):
    """Star 2 filling its Roche lobe (roche2 = 1): every element on
    the equipotential of L1, normals along the potential gradient, stored
    gravities = |grad Phi| / |grad Phi| at the back point (grid.rs:166-181),
    log g = log10(G M / a^2 * flux-weighted mean |grad Phi|), and the
    volume-averaged radius (comp_radius2) against a direct quadrature of
    the lobe's volume and Eggleton's formula (1/q: star 2's mass over star
    1's is q)."""
    pot, grad, xl1, mu = _roche(q)
    vs, tp = 400.0, 0.15
    # r2 is the lobe radius itself (to 1e-12), not -1: for r2 < 0 the
    # reference's comp_gravity2 (flux.rs:318-349) keeps the negative value in
    # its filling factor, where the grid uses the lobe radius, so its log g
    # then refers to another surface; the engine keeps that quirk.
    txt = W.model_text({"q": q, "r2": (1.0 - xl1) * (1 - 1e-12), "velocity_scale": vs, "tperiod": tp,
                        "nlat2f": 60, "nlat2c": 60, "nlat1f": 10, "nlat1c": 10, **NO_ECL})
    _, sc = pyoracle.light_curve(txt, [0.2, 0.3])
    ob = pyoracle.OracleBatch(txt, [0.0], 0.001, 1)
    g = ob.grid(0, 2)
    pl1 = pot(np.array([xl1, 0.0, 0.0]))
    gr = grad(g["posn"])
    gn = np.linalg.norm(gr, axis=1)
    # surface positions to the chop accuracy delta_phase / 10 = 1e-8 in radius
    assert np.max(np.abs(pot(g["posn"]) - pl1) / np.maximum(gn, 1e-3)) < 2e-8
    ok = gn > 1e-3 * gn.max()    # the L1 nose has no defined normal
    assert np.allclose(g["dirn"][ok], gr[ok] / gn[ok, None], atol=1e-6)
    xl2 = brentq(lambda x: grad(np.array([x, 0.0, 0.0]))[0], 1.0 + 1e-3, 2.0, xtol=1e-14)
    xb = brentq(lambda x: pot(np.array([x, 0.0, 0.0])) - pl1, 1.0 + 1e-3, xl2, xtol=1e-14)
    gback = np.linalg.norm(grad(np.array([xb, 0.0, 0.0])))
    assert np.allclose(ob.gravity(0, 2), gn / gback, rtol=1e-6, atol=1e-7)
    a_m, gm = kepler(vs, tp)
    gmean = np.sum(g["flux"] * gn) / np.sum(g["flux"])
    assert abs(sc["logg2"] - math.log10(100.0 * gm / a_m ** 2 * gmean)) < 2e-6
    rv = _lobe_volume_radius(q)
    assert abs(sc["rv2"] / rv - 1) < 2e-3, (sc["rv2"], rv)
    assert abs(rv / eggleton(q) - 1) < 0.01


# ---------------------------- star spots and the equatorial band (brightness.rs:13-147)
def _element_temperatures(txt, grid):
    ob = pyoracle.OracleBatch(txt, [0.0], 0.001, 1)
    g = ob.grid(0, grid)
    return g, brightness_temperature(g["flux"] / g["area"])


def test_gaussian_star_spots():
    """A star spot of peak tcen and full width at half maximum fwhm (deg):
    T = t + (tcen - t) 2^(-(2 theta / fwhm)^2), theta the angle from the
    spot centre seen from the star's centre; star 1 and star 2 (measured
    from star 2's centre). No gravity darkening or irradiation."""
    spots = {"stsp11_long": 30.0, "stsp11_lat": 20.0, "stsp11_fwhm": 40.0, "stsp11_tcen": 9000.0,
             "stsp21_long": 200.0, "stsp21_lat": -10.0, "stsp21_fwhm": 30.0, "stsp21_tcen": 6000.0}
    txt = W.model_text(spots, sphere_model(nlat=60))
    for grid, cen, t0, key in ((0, np.zeros(3), 15000.0, "stsp11"), (2, np.array([1.0, 0, 0]), 4000.0, "stsp21")):
        g, temp = _element_temperatures(txt, grid)
        lo, la = math.radians(spots[key + "_long"]), math.radians(spots[key + "_lat"])
        sdir = np.array([math.cos(la) * math.cos(lo), math.cos(la) * math.sin(lo), math.sin(la)])
        off = g["posn"] - cen
        th = np.degrees(np.arccos(np.clip(off @ sdir / np.linalg.norm(off, axis=1), -1, 1)))
        expect = t0 + (spots[key + "_tcen"] - t0) * 2.0 ** (-(2 * th / spots[key + "_fwhm"]) ** 2)
        assert np.allclose(temp, expect, rtol=2e-6, atol=0), (grid, np.max(np.abs(temp / expect - 1)))
        assert np.max(temp) > t0 + 0.8 * (spots[key + "_tcen"] - t0) or np.min(temp) < t0 - 0.8 * (t0 - spots[key + "_tcen"])


def test_uniform_equatorial_spot():
    """The equatorial band (uesp_*): uesp_temp inside |latitude| <= lathw and
    the longitude range; outside, the excess over t1 falls by exp(-x /
    taper) for each degree x beyond the band in latitude and in longitude."""
    band = {"uesp_long1": 20.0, "uesp_long2": 80.0, "uesp_lathw": 15.0, "uesp_taper": 10.0,
            "uesp_temp": 10000.0}
    t1 = 15000.0
    txt = W.model_text(band, sphere_model(nlat=60))
    g, temp = _element_temperatures(txt, 0)
    p = g["posn"]
    lat = np.degrees(np.arcsin(p[:, 2] / np.linalg.norm(p, axis=1)))
    lng = np.degrees(np.arctan2(p[:, 1], p[:, 0]))
    dlong = np.abs((lng - 50.0 + 180.0) % 360.0 - 180.0)
    flat = np.exp(-np.maximum(np.abs(lat) - 15.0, 0.0) / 10.0)
    flong = np.exp(-np.maximum(dlong - 30.0, 0.0) / 10.0)
    expect = t1 + (10000.0 - t1) * flat * flong
    inside = (np.abs(lat) <= 15.0) & (dlong <= 30.0)
    assert inside.sum() > 20
    assert np.allclose(temp, expect, rtol=2e-6, atol=0), np.max(np.abs(temp / expect - 1))
