"""GPU parity: the HIP engine (through the C ABI) against the FP64 oracle.

Bar: max relative flux error <= 1e-6 against the FP64 oracle (FP32 element
terms summed per patch in FP32 and across patches in FP64, FP64 patch
moments; the observed level is ~1e-7), and eclipse
visibility masks identical bit for bit (the FP64 interval test of
types.rs:50-58). Chi-squared is compared at 1e-5 relative (it amplifies
flux errors by the signal-to-noise ratio).
"""
import os

import numpy as np
import pytest

import pyoracle
from conftest import GOLDEN
from lcurve_rs import workloads as W
from variants import ISCALE_DISC, VARIANTS

pytestmark = pytest.mark.gpu

FLUX_TOL = 1e-6
CHISQ_TOL = 1e-5


def gpu_curve(lc, text, times, expose, ndiv, **kw):
    m = lc.Model.from_text(text)
    return m.light_curve(times=times, expose=expose, ndiv=ndiv, **kw)


def max_rel(a, b):
    return float(np.max(np.abs(a - b) / np.abs(b)))


def check_curve(lc, text, times, expose=0.001, ndiv=1, tol=FLUX_TOL):
    g = gpu_curve(lc, text, times, expose, ndiv)
    r, sc = pyoracle.light_curve(text, times, expose=expose, ndiv=ndiv, nthreads=8)
    err = max_rel(g.calc, r)
    assert err <= tol, err
    assert abs(g.wdwarf / sc["wdwarf"] - 1) <= tol
    for k in ("logg1", "logg2", "rv1", "rv2"):
        assert getattr(g, k) == pytest.approx(sc[k], rel=1e-12, abs=1e-300)
    return err


def check_masks(lc, text, phases, grids=range(7)):
    b = lc.Batch(lc.Model.from_text(text), phases, 0.001, 1, keep_maps=True)
    ob = pyoracle.OracleBatch(text, phases, 0.001, 1)
    total = 0
    for g in grids:
        gm = b.visibility(0, g, phases)
        om = ob.visibility(0, g, phases)
        assert gm.shape == om.shape, (g, gm.shape, om.shape)
        assert np.array_equal(gm, om), (g, int((gm != om).sum()))
        total += gm.size
    return total


# ------------------------------------------------------------ golden curve
def test_golden_curve(lcurve_rs, golden_model_text, golden_lc):
    m = lcurve_rs.Model(os.path.join(GOLDEN, "test_model.dat"))
    res = m.light_curve(time1=-0.2, time2=1.2, ntime=500, expose=0.001, ndiv=1)
    # tests/regression.rs:84-88 asserts < 1e-6 against the C++ lroche.
    assert max_rel(res.calc, golden_lc[:, 2]) < FLUX_TOL
    check_curve(lcurve_rs, golden_model_text, golden_lc[:, 0])


def test_golden_masks(lcurve_rs, golden_model_text):
    ph = np.linspace(-0.3, 1.3, 1601)
    assert check_masks(lcurve_rs, golden_model_text, ph, grids=range(4)) > 0


# ------------------------------------------------------------ model features
@pytest.mark.parametrize("name", sorted(VARIANTS))
def test_feature_variant(lcurve_rs, golden_model_text, name):
    text = W.model_text(VARIANTS[name], golden_model_text)
    times = W.linspace_times(-0.2, 1.2, 300)
    check_curve(lcurve_rs, text, times)


DENSE_VARIANTS = ["base", "mucrit", "beaming", "lensing", "claret_ld", "poly4_ld", "disc_edge_spot",
                  "opaque_disc_calibration", "roche_primary"]


@pytest.mark.parametrize("name", DENSE_VARIANTS)
def test_dense_times_variant(lcurve_rs, golden_model_text, name):
    """2000 times over 1.4 orbits: 64-lane tiles span ~16 degrees of phase, so
    the kernel skips back-facing patches, sums whole patches from their
    moments and drops eclipse tests where no lane is eclipsed."""
    text = golden_model_text if name == "base" else W.model_text(VARIANTS[name], golden_model_text)
    check_curve(lcurve_rs, text, W.linspace_times(-0.2, 1.2, 2000))


@pytest.mark.parametrize("name", ["opaque_disc_calibration", "disc_edge_spot", "roche_primary"])
def test_feature_variant_masks(lcurve_rs, golden_model_text, name):
    text = W.model_text(VARIANTS[name], golden_model_text)
    check_masks(lcurve_rs, text, np.linspace(-0.25, 1.25, 601))


def test_exposure_smearing(lcurve_rs, golden_model_text):
    times = W.linspace_times(-0.2, 1.2, 200)
    check_curve(lcurve_rs, golden_model_text, times, expose=0.01, ndiv=5)
    check_curve(lcurve_rs, golden_model_text, times, expose=0.02, ndiv=2)


def test_mixed_ndiv_per_point(lcurve_rs, golden_model_text, tmp_path):
    t = W.linspace_times(-0.1, 1.1, 150)
    nd = (np.arange(150) % 4) + 1
    f = tmp_path / "data.dat"
    with open(f, "w") as fh:
        for i in range(150):
            fh.write(f"{float(t[i])!r} 0.004 1e-9 1e-11 1 {int(nd[i])}\n")
    m = lcurve_rs.Model.from_text(golden_model_text)
    g = m.light_curve(data=str(f))
    r, sc = pyoracle.light_curve(golden_model_text, t, expose=0.004, ndiv=nd, flux=np.full(150, 1e-9),
                                 ferr=np.full(150, 1e-11), weight=np.ones(150), rdata=True)
    assert max_rel(g.calc, r) <= FLUX_TOL
    assert g.chisq == pytest.approx(sc["chisq"], rel=CHISQ_TOL)
    assert g.wnok == sc["wnok"]


# ------------------------------------------------------------ workloads
def test_wd_donor_45k_full(lcurve_rs):
    text, times, expose, ndiv = W.wd_donor()
    check_curve(lcurve_rs, text, times, expose, ndiv)
    check_masks(lcurve_rs, text, W.linspace_times(-0.2, 1.2, 500), grids=range(4))


def test_disc_spot_reduced(lcurve_rs):
    text, times, expose, ndiv = W.disc_spot(scale=0.3, ntime=400)
    check_curve(lcurve_rs, text, times, expose, ndiv)
    check_masks(lcurve_rs, text, np.linspace(-0.5, 1.5, 301))


def test_disc_spot_full_sampled(lcurve_rs):
    """Full 300k-element configuration: the GPU curve over all 5000 times
    against the oracle at 40 sampled times (the polynomial factor is 1 for
    this model, so a subset of times gives the same fluxes)."""
    text, times, expose, ndiv = W.disc_spot()
    g = gpu_curve(lcurve_rs, text, times, expose, ndiv)
    idx = np.linspace(0, len(times) - 1, 40).astype(int)
    r, _ = pyoracle.light_curve(text, times[idx], expose=expose, ndiv=ndiv, nthreads=16)
    assert max_rel(g.calc[idx], r) <= FLUX_TOL


# ------------------------------------------------------------ batched chi^2
def batch_data(n=200, flux=1e-9, ferr=1e-11):
    t = W.linspace_times(-0.2, 1.2, n)
    return t, np.full(n, flux), np.full(n, ferr), np.ones(n)


@pytest.mark.parametrize("scale", [False, True])
def test_chisq_batch_matches_oracle(lcurve_rs, golden_model_text, tmp_path, scale):
    # crates/lcurve-cuda/tests/gpu_correctness.rs:28-157 with the tolerance
    # tightened from 1e-3 to 1e-5.
    t, fl, fe, wt = batch_data()
    f = tmp_path / "d.dat"
    np.savetxt(f, np.column_stack([t, np.full(len(t), 0.001), fl, fe, wt, np.ones(len(t))]), fmt="%.17g")
    frac = np.linspace(-0.5, 0.5, 64)
    vals = np.column_stack([0.5 + 0.02 * frac, 82.0 + 2.0 * frac])
    m = lcurve_rs.Model.from_text(golden_model_text)
    g = m.chisq_batch(str(f), ["q", "iangle"], vals, scale=scale)
    ob = pyoracle.OracleBatch(golden_model_text, t, 0.001, 1, ["q", "iangle"], vals, flux=fl, ferr=fe,
                              weight=wt, nthreads=8)
    r, _ = ob.flux(scale=scale, nthreads=8)
    assert np.all(np.isfinite(g)) and np.all(np.isfinite(r))
    assert np.max(np.abs(g / r - 1)) < CHISQ_TOL


def test_chisq_batch_unphysical_set_is_nan(lcurve_rs, golden_model_text, tmp_path):
    t, fl, fe, wt = batch_data(50)
    f = tmp_path / "d.dat"
    np.savetxt(f, np.column_stack([t, np.full(len(t), 0.001), fl, fe, wt, np.ones(len(t))]), fmt="%.17g")
    m = lcurve_rs.Model.from_text(golden_model_text)
    g = m.chisq_batch(str(f), ["r2"], np.array([[0.99], [-1.0], [0.5]]), scale=False)
    assert np.isnan(g[0]) and np.isfinite(g[1])
    ob = pyoracle.OracleBatch(golden_model_text, t, 0.001, 1, ["r2"], [[0.99], [-1.0], [0.5]], flux=fl,
                              ferr=fe, weight=wt)
    r, _ = ob.flux(scale=False)
    assert np.isnan(r[0]) == np.isnan(g[0]) and np.isnan(r[2]) == np.isnan(g[2])


def uneven_batch(n_sets=12, ntime=300, seed=7):
    sets = W.chisq_sets(n_sets, seed=seed, nlat_lo=30, nlat_hi=90)
    base = W.chisq_base_text()
    texts = [W.model_text({"q": s[0], "iangle": s[1], "r1": s[2], "t1": s[3], "t2": s[4],
                           "nlat1f": 24, "nlat1c": 24, "nlat2f": int(s[5]), "nlat2c": int(s[5])}, base)
             for s in sets]
    return texts, W.linspace_times(-0.2, 1.2, ntime)


def test_uneven_batch_per_set_observations(lcurve_rs):
    texts, t = uneven_batch()
    models = [lcurve_rs.Model.from_text(x) for x in texts]
    b = lcurve_rs.Batch(models, t, 0.001, 1)
    true = b.run(scale=False)["flux"]
    noise = W.chisq_noise(len(texts), len(t), seed=11)
    obs = true * (1 + 0.01 * noise)
    sig = 0.01 * np.abs(true)
    b.set_obs(obs, sig, np.ones_like(obs))
    out = b.run(scale=True)
    ob = pyoracle.OracleBatch(texts, t, 0.001, 1, flux2d=obs, ferr2d=sig, weight2d=np.ones_like(obs),
                              nthreads=8)
    r, rf = ob.flux(scale=True, nthreads=8)
    assert np.max(np.abs(out["flux"] / rf - 1)) <= FLUX_TOL
    assert np.max(np.abs(out["chisq"] / r - 1)) < CHISQ_TOL
    # chi^2 per point ~ 1 for 1% noise scored against its own model
    assert 0.8 < np.median(out["chisq"] / len(t)) < 1.2


def test_graph_replay_is_stable(lcurve_rs):
    texts, t = uneven_batch(6, 200)
    b = lcurve_rs.Batch([lcurve_rs.Model.from_text(x) for x in texts], t, 0.001, 1)
    first = b.run(scale=True)
    for _ in range(4):
        lcurve_rs.Model.from_text(texts[0]).light_curve(ntime=10)  # other work in between
        again = b.run(scale=True)
        assert np.array_equal(first["flux"], again["flux"])
        assert np.array_equal(first["chisq"], again["chisq"])


def _chisq_texts(n, nlat_lo=89, nlat_hi=344, seed=42):
    base = W.chisq_base_text()
    return [W.model_text({"q": s[0], "iangle": s[1], "r1": s[2], "t1": s[3], "t2": s[4],
                          "nlat2f": int(s[5]), "nlat2c": int(s[5])}, base)
            for s in W.chisq_sets(n, seed=seed, nlat_lo=nlat_lo, nlat_hi=nlat_hi)]


def test_limb_lanes_match_full_tiles_and_oracle(lcurve_rs):
    """No-eclipse patches split per lane (moments / limb lanes regrouped
    across tiles / nothing) against the same batch with every lane of the
    tile running the element loop, and against the FP64 oracle."""
    texts = _chisq_texts(6)
    t = W.linspace_times(-0.2, 1.2, 2000)
    models = [lcurve_rs.Model.from_text(x) for x in texts]
    try:
        lcurve_rs.set_limb(False)
        off = lcurve_rs.Batch(models, t, 0.001, 1)
        f_off = off.run(scale=False)["flux"]
        assert off.stats()["limb_chunks"] == 0
    finally:
        lcurve_rs.set_limb(True)
    on = lcurve_rs.Batch(models, t, 0.001, 1)
    st = on.stats()
    assert st["limb_chunks"] > 0 and 0 < st["pairs_limb"] < st["pairs_evaluated"]
    assert st["pairs_evaluated"] < 0.8 * off.stats()["pairs_evaluated"]
    f_on = on.run(scale=False)["flux"]
    assert np.max(np.abs(f_on / f_off - 1)) <= 3e-7
    idx = np.linspace(0, len(t) - 1, 60).astype(int)
    ob = pyoracle.OracleBatch(texts, t[idx], 0.001, 1, nthreads=16)
    _, rf = ob.flux(scale=False, nthreads=16)
    assert np.max(np.abs(f_on[:, idx] / rf - 1)) <= FLUX_TOL
    # fixed summation order: replays are identical
    again = on.run(scale=False)["flux"]
    assert np.array_equal(f_on, again)


@pytest.mark.parametrize("max_part", [16, 64])
def test_limb_thresholds_and_split_items(lcurve_rs, max_part):
    """Other limb thresholds give other mixes of whole and split limb items
    (a patch's remainder of <= 32 or <= 16 limb lanes shares its elements
    among 2 or 4 lane groups): every mix matches the tiles without limb
    lanes, and the lane fill stays within [0, 1]."""
    texts = _chisq_texts(4)
    t = W.linspace_times(-0.2, 1.2, 1500)
    models = [lcurve_rs.Model.from_text(x) for x in texts]
    try:
        lcurve_rs.set_limb(False)
        f_off = lcurve_rs.Batch(models, t, 0.001, 1).run(scale=False)["flux"]
        lcurve_rs.set_limb(True, max_part)
        b = lcurve_rs.Batch(models, t, 0.001, 1)
        st = b.stats()
        f_on = b.run(scale=False)["flux"]
    finally:
        lcurve_rs.set_limb(True)
    assert st["limb_chunks"] > 0 and 0 < st["pairs_limb"] <= st["limb_slots"]
    # Split items take fewer slots than one whole item per patch would.
    assert st["limb_slots"] < st["limb_items"] * 64 * 64
    assert np.max(np.abs(f_on / f_off - 1)) <= 3e-7
    assert np.array_equal(f_on, b.run(scale=False)["flux"])


def test_launch_forms_and_autotune_agree(lcurve_rs):
    texts, t = uneven_batch(6, 200)
    models = [lcurve_rs.Model.from_text(x) for x in texts]
    b = lcurve_rs.Batch(models, t, 0.001, 1)
    true = b.run(scale=False)["flux"]
    obs = 1.01 * true * (1 + 0.01 * W.chisq_noise(len(texts), len(t), seed=3))
    b.set_obs(obs, 0.01 * true, np.ones_like(true))
    ref = b.run(scale=True)
    b.autotune(scale=True)
    tuned = b.run(scale=True)
    st = b.stats()
    assert st["chunk"] > 0
    assert np.max(np.abs(b.run(scale=False)["flux"] / true - 1)) < 1e-6
    assert np.max(np.abs(tuned["flux"] / ref["flux"] - 1)) < 1e-6
    assert np.max(np.abs(tuned["chisq"] / ref["chisq"] - 1)) < 1e-5
    assert np.allclose(ref["sfac"], 1.01, rtol=3e-3)


def test_light_curve_result_fields(lcurve_rs, golden_model_text):
    m = lcurve_rs.Model.from_text(golden_model_text)
    res = m.light_curve(time1=0.0, time2=1.0, ntime=11)
    assert len(res.calc) == 11 and np.array_equal(res.calc, res.flux)
    assert np.allclose(res.times, np.linspace(0, 1, 11))
    assert list(res.sfac) == [1.0] * 5
    assert res.chisq == 0.0 and res.wnok == 0.0
    assert "LcResult(npoints=11" in repr(res)


def test_light_curve_scale_with_data(lcurve_rs, golden_model_text, tmp_path):
    t = W.linspace_times(-0.2, 1.2, 120)
    r, _ = pyoracle.light_curve(golden_model_text, t)
    obs = 3.0 * r
    f = tmp_path / "d.dat"
    np.savetxt(f, np.column_stack([t, np.full(len(t), 0.001), obs, 0.01 * obs, np.ones(len(t)),
                                   np.ones(len(t))]), fmt="%.17g")
    res = lcurve_rs.Model.from_text(golden_model_text).light_curve(data=str(f), scale=True)
    assert res.sfac[0] == pytest.approx(3.0, rel=1e-6)
    assert np.max(np.abs(res.calc / obs - 1)) < 1e-6
    assert res.wnok == len(t)


# ------------------------------------------------------------ iscale
def _iscale_obs(text, t, truth, seed):
    fcomp, ncol = pyoracle.components(text, t, nthreads=8)
    clean = fcomp @ np.asarray(truth)
    ferr = 0.003 * np.abs(clean)
    y = clean + ferr * np.random.default_rng(seed).standard_normal(len(t))
    w = np.ones(len(t))
    w[::9] = 0.0
    return y, ferr, w


def _iscale_check(lc, text, t, expose=0.001, ndiv=1, scale=True, truth=(1.3, 0.7, 2.0, 0.5, 0.9)):
    y, ferr, w = _iscale_obs(text, t, truth, 11)
    r, sc = pyoracle.light_curve(text, t, expose=expose, ndiv=ndiv, flux=y, ferr=ferr, weight=w,
                                 scale=scale, rdata=True, nthreads=8)
    nd = np.full(len(t), ndiv, dtype=np.int32)
    m = lc.Model.from_text(text)
    calc = np.zeros(len(t))
    res = lc._native.LcResultC()
    lc._native.check(lc._native.lib().lch_light_curve(
        lc._ctx(), m._h, lc._dptr(np.ascontiguousarray(t, dtype=np.float64)),
        lc._dptr(np.full(len(t), expose)), lc._dptr(y), lc._dptr(ferr), lc._dptr(w), lc._iptr(nd),
        len(t), int(scale), 1, lc._dptr(calc), lc.ctypes.byref(res)))
    assert max_rel(calc, r) <= FLUX_TOL
    assert res.chisq == pytest.approx(sc["chisq"], rel=CHISQ_TOL)
    assert res.wnok == sc["wnok"]
    assert np.allclose(list(res.sfac), sc["sfac"], rtol=1e-5, atol=1e-12)
    assert res.wdwarf == pytest.approx(sc["wdwarf"], rel=FLUX_TOL)
    return res, sc


@pytest.mark.parametrize("scale", [True, False])
def test_iscale_components_and_svd_factors(lcurve_rs, golden_model_text, scale):
    """Per-component columns and their SVD least-squares factors
    (orchestration.rs:181-297, 405-436) against the oracle, whose SVD is
    pinned to numpy.linalg.lstsq."""
    text = W.model_text(dict(ISCALE_DISC, iscale=1), golden_model_text)
    res, sc = _iscale_check(lcurve_rs, text, W.linspace_times(-0.2, 1.2, 240), scale=scale)
    if scale:
        assert np.allclose(sc["sfac"], [1.3, 0.7, 2.0, 0.5, 0.9], rtol=0.05)
    else:
        assert list(res.sfac) == [1.0] * 5


def test_iscale_smeared_exposures(lcurve_rs, golden_model_text):
    """ndiv > 1: comp_star1 weights by scale1(phase), comp_star2 by scale2(phi)."""
    text = W.model_text(dict(ISCALE_DISC, iscale=1), golden_model_text)
    _iscale_check(lcurve_rs, text, W.linspace_times(-0.2, 1.2, 120), expose=0.01, ndiv=5)


def test_iscale_absent_components(lcurve_rs, golden_model_text):
    """No disc, edge or spot: zero columns get factor exactly 0."""
    text = W.model_text({"iscale": 1}, golden_model_text)
    res, _ = _iscale_check(lcurve_rs, text, W.linspace_times(-0.2, 1.2, 200), truth=(1.1, 0, 0, 0, 0.8))
    assert list(res.sfac)[1:4] == [0.0, 0.0, 0.0]


def test_iscale_in_chisq_batch(lcurve_rs, golden_model_text):
    """Mixed batch: iscale sets fit per-component factors, the others autoscale."""
    t = W.linspace_times(-0.2, 1.2, 160)
    base = W.model_text(ISCALE_DISC, golden_model_text)
    texts = [W.model_text({"iscale": k % 2, "q": 0.1 + 0.02 * k}, base) for k in range(6)]
    y, ferr, w = _iscale_obs(W.model_text({"iscale": 1}, base), t, (1.2, 0.8, 1.5, 0.6, 1.1), 3)
    b = lcurve_rs.Batch([lcurve_rs.Model.from_text(x) for x in texts], t, 0.001, 1)
    b.set_obs(np.tile(y, (6, 1)), np.tile(ferr, (6, 1)), np.tile(w, (6, 1)))
    out = b.run(scale=True)
    ob = pyoracle.OracleBatch(texts, t, 0.001, 1, flux2d=np.tile(y, (6, 1)), ferr2d=np.tile(ferr, (6, 1)),
                              weight2d=np.tile(w, (6, 1)), nthreads=8)
    chi, flux = ob.flux(scale=True, nthreads=8)
    assert np.allclose(out["chisq"], chi, rtol=CHISQ_TOL)
    assert np.max(np.abs(out["flux"] - flux) / np.abs(flux)) <= FLUX_TOL
    for k in range(6):
        if k % 2 == 0:
            assert np.all(out["sfac5"][k] == out["sfac"][k])
        else:
            assert len(set(out["sfac5"][k])) > 1


def test_setup_error_raises(lcurve_rs, golden_model_text):
    m = lcurve_rs.Model.from_text(W.model_text({"r2": 0.99}, golden_model_text))
    with pytest.raises(ValueError, match="Roche lobe"):
        m.light_curve(ntime=10)
