"""CPU tests of the FP64 oracle (and the shared CPU setup it runs on).

What pins the oracle (each test names the mistake it would catch):
  * the reference's golden curve, test_data/test_lc.dat from the C++ lroche
    (tests/regression.rs:9-88): Roche donor, spherical WD, gravity
    darkening, irradiation, both eclipses, fine/coarse Ginterp switching;
  * closed forms on spheres: projected area times surface brightness for no,
    linear, 4-term polynomial and Claret limb darkening; surface area 4 pi r^2;
  * brute-force ray tests of the eclipse intervals (sphere, Roche lobe, disc);
  * the reference's batch invariants (tests/chisq_batch.rs): batch equals
    sequential, NaN for an unphysical set;
  * re_scale / compute_chisq / smearing / polynomial-factor identities.
"""
import math

import numpy as np
import pytest

import pyoracle
from lcurve_rs import workloads as W

H, C, K = 6.6262e-34, 2.99792458e8, 1.3806e-23  # constants of crates/subs/src/constants.rs


def planck_nu(wave_nm, temp):
    """Planck's law B_nu in W/m^2/Hz/sr at wavelength wave_nm."""
    lam = wave_nm * 1e-9
    return 2.0 * H * C / lam ** 3 / math.expm1(H * C / (lam * K * temp))


# ------------------------------------------------------------------ golden
def test_oracle_matches_cpp_golden_curve(golden_model_text, golden_lc):
    times = W.linspace_times(-0.2, 1.2, 500)
    assert np.array_equal(times, golden_lc[:, 0])
    calc, sc = pyoracle.light_curve(golden_model_text, times, expose=0.001, ndiv=1, nthreads=4)
    rel = np.abs(calc - golden_lc[:, 2]) / np.abs(golden_lc[:, 2])
    # tests/regression.rs:84-88 asserts < 1e-6; the golden file has 10
    # significant digits, so agreement at ~1e-10 is the best possible.
    assert rel.max() < 1e-9, rel.max()


def test_golden_fails_if_ginterp_quirk_dropped(golden_model_text, golden_lc):
    """With equal fine and coarse grids the Ginterp scale factors do not
    enter; the golden curve then differs visibly, showing that the golden
    test exercises the calibration path."""
    times = W.linspace_times(-0.2, 1.2, 500)
    txt = W.model_text({"nlat1c": 20, "nlat2c": 40}, golden_model_text)
    calc, _ = pyoracle.light_curve(txt, times, expose=0.001, ndiv=1, nthreads=4)
    rel = np.abs(calc - golden_lc[:, 2]) / np.abs(golden_lc[:, 2])
    assert rel.max() > 1e-6


# ------------------------------------------------------------ closed forms
def sphere_model(ld=None, limb="Poly", nlat=80, r1=0.05, r2=0.2, t1=15000.0, t2=4000.0, iangle=60.0):
    ov = {"roche1": 0, "roche2": 0, "r1": r1, "r2": r2, "t1": t1, "t2": t2, "absorb": 0.0,
          "eclipse1": 0, "eclipse2": 0, "nlat1f": nlat, "nlat1c": nlat, "nlat2f": nlat, "nlat2c": nlat,
          "iangle": iangle, "limb1": limb, "limb2": limb, "gravity_dark1": 0.0, "gravity_dark2": 0.0}
    for i in range(4):
        ov[f"ldc1_{i + 1}"] = 0.0
        ov[f"ldc2_{i + 1}"] = 0.0
    for k, v in (ld or {}).items():
        ov[k] = v
    return W.model_text(ov)


def sphere_flux(r, temp, lint):
    """2 pi r^2 B int_0^1 I(mu) mu dmu with lint = int_0^1 I(mu) mu dmu."""
    return 2.0 * math.pi * r * r * planck_nu(550.0, temp) * lint


@pytest.mark.parametrize("nlat", [40, 80])
def test_uniform_spheres_give_projected_area_times_brightness(nlat):
    txt = sphere_model(nlat=nlat)
    calc, _ = pyoracle.light_curve(txt, np.linspace(0.0, 1.0, 7), nthreads=4)
    expect = sphere_flux(0.05, 15000.0, 0.5) + sphere_flux(0.2, 4000.0, 0.5)
    err = np.abs(calc / expect - 1.0)
    assert err.max() < 3.0 / nlat ** 2, err.max()


def test_sphere_discretisation_error_converges():
    errs = []
    for nlat in (30, 60, 120):
        calc, _ = pyoracle.light_curve(sphere_model(nlat=nlat), [0.3, 0.4], nthreads=4)
        expect = sphere_flux(0.05, 15000.0, 0.5) + sphere_flux(0.2, 4000.0, 0.5)
        errs.append(abs(calc[0] / expect - 1.0))
    assert errs[1] < errs[0] / 2 and errs[2] < errs[1] / 2, errs


@pytest.mark.parametrize("coeffs", [[0.4, 0, 0, 0], [0.3, 0.2, 0, 0], [0.5, -0.2, 0.3, -0.1]])
def test_polynomial_limb_darkening_closed_form(coeffs):
    # I = 1 - sum a_k (1-mu)^k;  int (1-mu)^k mu dmu = 1 / ((k+1)(k+2))
    ld = {f"ldc{s}_{k + 1}": coeffs[k] for s in (1, 2) for k in range(4)}
    txt = sphere_model(ld=ld, nlat=120)
    calc, _ = pyoracle.light_curve(txt, [0.1, 0.2], nthreads=4)
    lint = 0.5 - sum(a / ((k + 1) * (k + 2)) for k, a in enumerate(coeffs, start=1))
    expect = sphere_flux(0.05, 15000.0, lint) + sphere_flux(0.2, 4000.0, lint)
    assert abs(calc[0] / expect - 1.0) < 3e-4


@pytest.mark.parametrize("coeffs", [[0.55, -0.35, 0.65, -0.25], [0.7, -0.45, 0.85, -0.35]])
def test_claret_limb_darkening_closed_form(coeffs):
    # I = 1 - sum a_k (1 - mu^(k/2));  int mu^(k/2+1) dmu = 2 / (k+4)
    ld = {f"ldc{s}_{k + 1}": coeffs[k] for s in (1, 2) for k in range(4)}
    txt = sphere_model(ld=ld, limb="Claret", nlat=120)
    calc, _ = pyoracle.light_curve(txt, [0.1, 0.2], nthreads=4)
    lint = 0.5 - sum(a * (0.5 - 2.0 / (k + 4)) for k, a in enumerate(coeffs, start=1))
    expect = sphere_flux(0.05, 15000.0, lint) + sphere_flux(0.2, 4000.0, lint)
    assert abs(calc[0] / expect - 1.0) < 3e-4


def test_sphere_grid_area_and_brightness_sums():
    # crates/lcurve/examples/diag.rs:31-55: sum area ~ 4 pi r^2 and
    # sum B*A = B * sum area for a uniform sphere.
    txt = sphere_model(nlat=100)
    ob = pyoracle.OracleBatch(txt, [0.0, 0.5], 0.001, 1)
    g = ob.grid(0, 0)
    assert abs(g["area"].sum() / (4 * math.pi * 0.05 ** 2) - 1) < 1e-3
    b = planck_nu(550.0, 15000.0)
    assert abs(g["flux"].sum() / (b * g["area"].sum()) - 1) < 1e-6
    r = np.linalg.norm(g["posn"], axis=1)
    assert np.allclose(r, 0.05, rtol=1e-12)


# -------------------------------------------------------- eclipse intervals
def earth(iangle_deg, phase):
    i = math.radians(iangle_deg)
    return np.stack([math.sin(i) * np.cos(2 * np.pi * phase), -math.sin(i) * np.sin(2 * np.pi * phase),
                     np.full(np.shape(phase), math.cos(i))], axis=-1)


def test_sphere_eclipse_matches_ray_test():
    """Star-1 elements behind a spherical star 2: ray-sphere intersection."""
    txt = sphere_model(nlat=24, iangle=85.0)
    txt = W.model_text({"eclipse1": 1, "eclipse2": 1}, txt)
    phases = np.linspace(-0.1, 0.1, 401)
    ob = pyoracle.OracleBatch(txt, [0.0], 0.001, 1)
    g = ob.grid(0, 0)
    mask = ob.visibility(0, 0, phases)
    e = earth(85.0, phases)                      # [P, 3]
    d = g["posn"][None, :, :] - np.array([1.0, 0, 0])  # [1, N, 3]
    b = np.einsum("pk,nk->pn", e, d[0])           # e . (p - c)
    disc = b * b - (np.sum(d[0] ** 2, axis=1)[None, :] - 0.2 ** 2)
    hit = (b < 0) & (disc > 0)
    clear = np.abs(disc) > 1e-9
    assert (hit.sum() > 0) and ((~hit).sum() > 0)
    assert np.array_equal(mask.astype(bool)[clear], ~hit[clear])


def _roche_potential(q, x, y, z):
    """Synchronous Roche potential with the donor at (1,0,0), in the
    rotating frame of the binary (textbook form, no reference code)."""
    mu = q / (1.0 + q)
    r1 = np.sqrt(x * x + y * y + z * z)
    r2 = np.sqrt((x - 1.0) ** 2 + y * y + z * z)
    return -(1 - mu) / r1 - mu / r2 - 0.5 * ((x - mu) ** 2 + y * y)


def test_roche_lobe_eclipse_matches_ray_march(golden_model_text):
    """White-dwarf elements eclipsed by the Roche-lobe-filling donor of the
    golden model: a phase is eclipsed iff the line of sight dips below the
    L1 potential somewhere along it (fine ray march)."""
    from scipy.optimize import brentq
    q = 0.5
    mu = q / (1 + q)
    dphidx = lambda x: (1 - mu) / x ** 2 - mu / (1 - x) ** 2 - (x - mu)
    xl1 = brentq(dphidx, 0.1, 0.9)
    pl1 = _roche_potential(q, xl1, 0.0, 0.0)
    phases = np.linspace(-0.08, 0.08, 161)
    ob = pyoracle.OracleBatch(golden_model_text, [0.0], 0.001, 1)
    g = ob.grid(0, 0)   # white dwarf, fine grid
    mask = ob.visibility(0, 0, phases).astype(bool)
    e = earth(82.0, phases)
    lam = np.linspace(0.0, 2.0, 4001)
    sel = np.arange(0, len(g["posn"]), 7)
    agree = total = 0
    for n in sel:
        p = g["posn"][n]
        pts = p[None, None, :] + lam[None, :, None] * e[:, None, :]      # [P, L, 3]
        pot = _roche_potential(q, pts[..., 0], pts[..., 1], pts[..., 2])
        # only the donor's lobe can eclipse: it lies on the donor's side of
        # L1 and within 1 - x_L1 of the donor's centre (the potential also
        # falls below the L1 value around the white dwarf and far outside
        # the binary, where the centrifugal term dominates)
        d2 = np.sqrt((pts[..., 0] - 1.0) ** 2 + pts[..., 1] ** 2 + pts[..., 2] ** 2)
        pot = np.where((pts[..., 0] > xl1) & (d2 <= 1.0 - xl1), pot, np.inf)
        margin = pot.min(axis=1) - pl1
        ecl = margin < 0
        clear = np.abs(margin) > 2e-4
        agree += np.sum(mask[clear, n] == ~ecl[clear])
        total += np.sum(clear)
    assert total > 1000 and agree == total, (agree, total)


def test_disc_eclipse_matches_ray_march():
    """Donor elements behind an opaque flared disc: ray-march against
    |z| <= h R^beta for rdisc1 <= R <= rdisc2."""
    txt = W.model_text({"add_disc": 1, "opaque": 1, "rdisc1": 0.1, "rdisc2": 0.3, "nrad": 20,
                        "height_disc": 0.1, "beta_disc": 1.5, "eclipse1": 0, "eclipse2": 0,
                        "nlat2f": 30, "nlat2c": 30,
                        "nlat1f": 12, "nlat1c": 12})
    phases = np.linspace(0.3, 0.7, 81)
    ob = pyoracle.OracleBatch(txt, [0.0], 0.001, 1)
    g = ob.grid(0, 2)
    # eclipse2 = 0 removes the white dwarf's eclipses of donor elements, so
    # the donor grid carries only the opaque-disc intervals.
    mask = ob.visibility(0, 2, phases).astype(bool)
    e = earth(82.0, phases)
    lam = np.linspace(0.0, 3.0, 6001)
    agree = total = 0
    for n in range(0, len(g["posn"]), 5):
        p = g["posn"][n]
        pts = p[None, None, :] + lam[None, :, None] * e[:, None, :]
        rxy = np.hypot(pts[..., 0], pts[..., 1])
        h = 0.1 * np.power(np.maximum(rxy, 1e-12), 1.5)
        inside = (rxy >= 0.1) & (rxy <= 0.3) & (np.abs(pts[..., 2]) <= h)
        ecl = inside[:, 1:].any(axis=1)
        # distance of the ray from the disc surface, to skip grazing rays
        gap = np.where((rxy >= 0.1) & (rxy <= 0.3), np.abs(np.abs(pts[..., 2]) - h), np.inf)
        rim = np.where(np.abs(pts[..., 2]) <= h + 2e-3,
                       np.minimum(np.abs(rxy - 0.1), np.abs(rxy - 0.3)), np.inf)
        clear = (gap[:, 1:].min(axis=1) > 2e-3) & (rim[:, 1:].min(axis=1) > 2e-3)
        agree += np.sum(mask[clear, n] == ~ecl[clear])
        total += np.sum(clear)
    assert total > 500 and agree == total, (agree, total)


# --------------------------------------------------------- batch invariants
def make_data(n, flux=0.0, ferr=1.0, weight=1.0):
    t = W.linspace_times(-0.2, 1.2, n)
    return t, np.full(n, flux), np.full(n, ferr), np.full(n, weight)


def test_batch_single_matches_sequential(golden_model_text):
    # tests/chisq_batch.rs:26-46
    t, fl, fe, wt = make_data(100)
    ob = pyoracle.OracleBatch(golden_model_text, t, 0.001, 1, ["q", "iangle"], [[0.5, 82.0]],
                              flux=fl, ferr=fe, weight=wt)
    chisq, _ = ob.flux(scale=False)
    _, sc = pyoracle.light_curve(golden_model_text, t, flux=fl, ferr=fe, weight=wt, rdata=True)
    assert abs(chisq[0] - sc["chisq"]) < 1e-10 * max(abs(sc["chisq"]), 1.0)


def test_batch_matches_n_sequential(golden_model_text):
    # tests/chisq_batch.rs:49-89
    t, fl, fe, wt = make_data(100)
    offs = [-0.02, -0.01, 0.0, 0.01, 0.02]
    vals = [[0.5 + d, 82.0] for d in offs]
    ob = pyoracle.OracleBatch(golden_model_text, t, 0.001, 1, ["q", "iangle"], vals, flux=fl, ferr=fe,
                              weight=wt, nthreads=4)
    chisq, _ = ob.flux(scale=False, nthreads=4)
    for d, c in zip(offs, chisq):
        txt = W.model_text({"q": 0.5 + d}, golden_model_text)
        _, sc = pyoracle.light_curve(txt, t, flux=fl, ferr=fe, weight=wt, rdata=True)
        assert abs(c - sc["chisq"]) < 1e-10 * max(abs(sc["chisq"]), 1.0)


def test_unphysical_set_gives_nan(golden_model_text):
    # tests/chisq_batch.rs:92-109
    t, fl, fe, wt = make_data(50)
    ob = pyoracle.OracleBatch(golden_model_text, t, 0.001, 1, ["r2"], [[0.99], [-1.0]], flux=fl, ferr=fe,
                              weight=wt)
    chisq, _ = ob.flux(scale=False)
    assert np.isnan(chisq[0]) and np.isfinite(chisq[1])


def test_rescale_recovers_exact_multiple(golden_model_text):
    t = W.linspace_times(-0.2, 1.2, 120)
    base, _ = pyoracle.light_curve(golden_model_text, t)
    obs = 2.5 * base
    w = np.ones(len(t))
    w[::7] = 0.0          # excluded points must not matter
    obs[::7] = 1e3
    calc, sc = pyoracle.light_curve(golden_model_text, t, flux=obs, ferr=0.01 * np.abs(obs), weight=w,
                                    scale=True, rdata=True)
    assert abs(sc["sfac"][0] / 2.5 - 1) < 1e-13
    assert sc["chisq"] < 1e-18
    assert abs(sc["wnok"] - w.sum()) == 0
    assert np.allclose(calc, 2.5 * base, rtol=1e-13)


def test_chisq_of_known_offsets(golden_model_text):
    t = W.linspace_times(-0.2, 1.2, 60)
    base, _ = pyoracle.light_curve(golden_model_text, t)
    sig = 1e-12 * np.ones(len(t))
    k = np.linspace(-2, 2, len(t))
    _, sc = pyoracle.light_curve(golden_model_text, t, flux=base + k * sig, ferr=sig,
                                 weight=np.ones(len(t)), rdata=True)
    assert abs(sc["chisq"] / np.sum(k * k) - 1) < 1e-9


def test_smearing_a_constant_curve_is_identity():
    # Face-on: the flux does not depend on phase, so exposure smearing with
    # trapezoid weights must return it unchanged.
    txt = W.model_text({"iangle": 0.0, "nlat1f": 20, "nlat1c": 20, "nlat2f": 30, "nlat2c": 30})
    t = np.linspace(0.0, 1.0, 9)
    f1, _ = pyoracle.light_curve(txt, t, expose=0.05, ndiv=1)
    f5, _ = pyoracle.light_curve(txt, t, expose=0.05, ndiv=5)
    assert np.allclose(f1, f1[0], rtol=1e-13)
    assert np.allclose(f5, f1, rtol=1e-12)


def test_polynomial_factor_and_third_light(golden_model_text):
    t = W.linspace_times(0.2, 0.9, 50)
    base, _ = pyoracle.light_curve(golden_model_text, t)
    txt = W.model_text({"slope": 0.1, "quad": -0.05, "cube": 0.02, "third": 3e-11}, golden_model_text)
    f, _ = pyoracle.light_curve(txt, t)
    mid, rng = (t.min() + t.max()) / 2, (t.max() - t.min()) / 2
    x = (t - mid) / rng
    expect = (1 + x * (0.1 + x * (-0.05 + x * 0.02))) * base + 3e-11
    assert np.allclose(f, expect, rtol=1e-13)


def test_phase_newton_with_pdot_inverts_ephemeris():
    # orchestration.rs:231-237: t = t0 + P phi + pdot phi^2 (deltat = 0)
    txt = W.model_text({"t0": 0.3, "period": 0.2, "pdot": 1e-6})
    phi_true = np.linspace(-3.0, 40.0, 50)
    t = 0.3 + phi_true * (0.2 + 1e-6 * phi_true)
    assert np.allclose(pyoracle.phases(txt, t), phi_true, atol=1e-10)


# ------------------------------------------------------------ iscale (SVD)
from variants import ISCALE_DISC  # noqa: E402


def _weighted_lstsq(fcomp, ncol, y, ferr, w):
    ok = w > 0
    sw = np.sqrt(w[ok]) / ferr[ok]
    A = fcomp[ok, :ncol] * sw[:, None]
    b = y[ok] * sw
    return np.linalg.lstsq(A, b, rcond=None)[0]


def test_iscale_svd_matches_numpy_lstsq(golden_model_text):
    """svd_scale (orchestration.rs:405-436) against LAPACK's SVD least squares:
    observations built from known per-component factors plus noise."""
    text = W.model_text(dict(ISCALE_DISC, iscale=1), golden_model_text)
    t = W.linspace_times(-0.2, 1.2, 240)
    fcomp, ncol = pyoracle.components(text, t, nthreads=4)
    assert ncol == 5 and np.all(np.abs(fcomp).sum(0) > 0)
    truth = np.array([1.3, 0.7, 2.0, 0.5, 0.9])
    rng = np.random.default_rng(7)
    clean = fcomp @ truth
    ferr = 0.003 * np.abs(clean)
    y = clean + ferr * rng.standard_normal(len(t))
    w = np.ones(len(t))
    w[::7] = 0.0  # masked points take no part
    calc, sc = pyoracle.light_curve(text, t, flux=y, ferr=ferr, weight=w, scale=True, nthreads=4)
    ref = _weighted_lstsq(fcomp, ncol, y, ferr, w)
    assert np.allclose(sc["sfac"], ref, rtol=1e-9, atol=0)
    assert np.allclose(calc, fcomp @ sc["sfac"], rtol=1e-12)
    r = (y - calc) / ferr
    assert sc["chisq"] == pytest.approx(np.sum(w * r * r), rel=1e-12)
    assert sc["wnok"] == pytest.approx(w.sum())


def test_iscale_absent_components_get_zero_factor(golden_model_text):
    """No disc, edge or spot: those columns are zero, their singular values
    are 0 and the minimum-norm solution gives them factor 0, as LAPACK does."""
    text = W.model_text({"iscale": 1}, golden_model_text)
    t = W.linspace_times(-0.2, 1.2, 200)
    fcomp, ncol = pyoracle.components(text, t, nthreads=4)
    assert np.all(fcomp[:, 1:4] == 0.0)
    y = fcomp @ np.array([1.1, 0.0, 0.0, 0.0, 0.8])
    ferr = 0.01 * y
    calc, sc = pyoracle.light_curve(text, t, flux=y, ferr=ferr, weight=np.ones(200), scale=True, nthreads=4)
    ref = _weighted_lstsq(fcomp, ncol, y, ferr, np.ones(200))
    assert np.allclose(sc["sfac"], ref, rtol=1e-9, atol=1e-12)
    assert sc["sfac"][0] == pytest.approx(1.1, rel=1e-9) and sc["sfac"][4] == pytest.approx(0.8, rel=1e-9)
    assert np.all(sc["sfac"][1:4] == 0.0)
    assert sc["chisq"] < 1e-12 * len(t)


def test_iscale_unscaled_columns_sum_to_the_light_curve(golden_model_text):
    """With ndiv = 1, no beaming and no third light the unscaled columns add
    up to comp_light; the only differences are the inclination's
    TWOPI*i/360 vs to_radians (last-bit) and Ginterp's scale1(phase) vs
    scale1(phi), equal when ndiv = 1."""
    text = W.model_text(ISCALE_DISC, golden_model_text)
    t = W.linspace_times(-0.2, 1.2, 150)
    plain, _ = pyoracle.light_curve(text, t, nthreads=4)
    fcomp, ncol = pyoracle.components(text, t, nthreads=4)
    assert np.allclose(fcomp.sum(1), plain, rtol=1e-12)
    text_i = W.model_text(dict(ISCALE_DISC, iscale=1), golden_model_text)
    calc, sc = pyoracle.light_curve(text_i, t, nthreads=4)
    assert np.array_equal(calc, fcomp[:, 0] + fcomp[:, 1] + fcomp[:, 2] + fcomp[:, 3] + fcomp[:, 4])
    assert np.all(sc["sfac"] == 1.0)


def test_iscale_too_few_points_gives_unit_factors(golden_model_text):
    text = W.model_text(dict(ISCALE_DISC, iscale=1), golden_model_text)
    t = W.linspace_times(-0.2, 1.2, 8)
    w = np.zeros(8)
    w[:4] = 1.0  # 4 usable points < 5 columns
    calc, sc = pyoracle.light_curve(text, t, flux=np.ones(8), ferr=np.ones(8), weight=w, scale=True)
    assert np.all(sc["sfac"] == 1.0)
