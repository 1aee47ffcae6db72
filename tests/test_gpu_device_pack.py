"""GPU checks of the device-packed batch (lch_dpack_*, lcurve_rs.DevicePackBatch).

For each model variant of the flux path, the per-set pack the GPU makes from
the raw grids must equal the host packer's bit for bit: element records in
patch order, patch cones and moments, work codes, segment flags (linear-law
rescaling), per-lane splits, work items with their lane-mask groups and moment
sums, and the limb tables (lch_dpack_check). Its curves and chi-squared values
must equal those of a host-packed Batch of the same models (the same chunk and
limb threshold) bit for bit, and through it the FP64 oracle within the parity
bar. A set whose setup fails gives NaN, as on the host path.
"""
import numpy as np
import pytest

import pyoracle
from lcurve_rs import workloads as W
from variants import ISCALE_DISC, VARIANTS

pytestmark = pytest.mark.gpu

CHECK_KEYS = ("record_bytes", "cones", "moments", "codes", "seg_flags", "items", "item_moments", "limb")


@pytest.fixture
def host_tuning(lcurve_rs):
    """Host-packed batches built in the test take the device pack's limb
    threshold; the context's defaults come back afterwards."""
    yield lcurve_rs
    lcurve_rs.set_limb(True)
    lcurve_rs.set_tuning(1024)


def _models(lc, text, n, seed):
    """n models around `text`: mass ratio and inclination varied."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        m = lc.Model.from_text(text)
        m.set_param("q", m.get_param("q") * (1.0 + 0.1 * (rng.random() - 0.5)))
        m.set_param("iangle", m.get_param("iangle") + 2.0 * (rng.random() - 0.5))
        out.append(m)
    return out


def _obs(n, seed):
    rng = np.random.default_rng(seed)
    return {"flux": 1.0 + 0.01 * rng.standard_normal(n), "ferr": np.full(n, 0.01), "weight": np.ones(n)}


def _host_limb(lc, limb):
    """Host-packed batches under the device pack's split rule (no
    element-level check of the cone-undecided lanes)."""
    lc.set_limb(limb > 0, max(limb, 1), check_elements=False)


def _compare(lc, models, times, expose, ndiv, chunk, obs, limb=0):
    dp = lc.DevicePackBatch(models, times, expose, ndiv, chunk=chunk, limb=limb, **obs)
    for s in range(len(models)):
        d = dp.check(s)
        assert all(d[k] == 0 for k in CHECK_KEYS), (s, d)
        assert d["host_items"] > 0, (s, d)
    rd = dp.run()
    _host_limb(lc, limb)
    lc.set_tuning(chunk)
    hb = lc.Batch(models, times, expose, ndiv, **obs)
    rh = hb.run()
    assert np.array_equal(rd["flux"], rh["flux"]), np.nanmax(np.abs(rd["flux"] / rh["flux"] - 1))
    for k in ("chisq", "wnok", "sfac", "wdwarf"):
        assert np.array_equal(rd[k], rh[k]), (k, rd[k], rh[k])
    return dp, rd


@pytest.mark.parametrize("limb", [0, 48])
@pytest.mark.parametrize("name", ["base"] + sorted(VARIANTS))
def test_device_pack_equals_host_pack(host_tuning, golden_model_text, name, limb):
    lc = host_tuning
    text = golden_model_text if name == "base" else W.model_text(VARIANTS[name], golden_model_text)
    times = W.linspace_times(-0.2, 1.2, 300)
    _compare(lc, _models(lc, text, 3, 1), times, 0.001, 1, 1024, _obs(len(times), 2), limb=limb)


@pytest.mark.parametrize("limb", [0, 16, 64])
def test_device_pack_smearing_small_chunks_and_iscale(host_tuning, golden_model_text, limb):
    """Exposure sub-phases, several chunks per segment, the disc, rim and
    spot, and the per-component fits; limb thresholds from few split pairs
    to every no-eclipse pair."""
    lc = host_tuning
    times = W.linspace_times(-0.3, 1.3, 257)
    text = W.model_text(VARIANTS["disc_edge_spot"], golden_model_text)
    _compare(lc, _models(lc, text, 2, 3), times, 0.004, 3, 256, _obs(len(times), 4), limb=limb)
    text = W.model_text(dict(ISCALE_DISC, iscale=1), golden_model_text)
    _compare(lc, _models(lc, text, 2, 5), times, 0.002, 2, 512, _obs(len(times), 6), limb=limb)


def test_device_pack_against_oracle(host_tuning, golden_model_text):
    """The curves of a device-packed batch against the FP64 oracle (bar:
    1e-6 relative), for a disc and bright-spot model with smearing."""
    lc = host_tuning
    text = W.model_text(VARIANTS["opaque_disc_calibration"], golden_model_text)
    times = W.linspace_times(-0.2, 1.2, 200)
    m = lc.Model.from_text(text)
    ref, _ = pyoracle.light_curve(text, times, expose=0.002, ndiv=3, nthreads=8)
    for limb in (0, 48):
        dp = lc.DevicePackBatch([m], times, 0.002, 3, chunk=2048, limb=limb)
        r = dp.run(scale=False)
        assert np.max(np.abs(r["flux"][0] / ref - 1)) <= 1e-6, limb


@pytest.mark.parametrize("limb", [0, 48])
def test_device_pack_many_sets_and_repeat(host_tuning, limb):
    """48 sets of the chi-squared workload's geometry at reduced grids, one
    of them unphysical: codes, items and limb tables equal the host's on a
    sample of sets, results equal the host-packed batch's, the failed set is
    NaN, and a re-upload and second launch give the same results."""
    lc = host_tuning
    base = W.chisq_base_text()
    p = W.chisq_sets(48, seed=7, nlat_lo=20, nlat_hi=60)
    models = []
    for i, s in enumerate(p):
        ov = {"q": s[0], "iangle": s[1], "r1": s[2], "t1": s[3], "t2": s[4], "nlat1f": 24, "nlat1c": 24,
              "nlat2f": int(s[5]), "nlat2c": int(s[5])}
        if i == 17:
            ov["r2"] = 0.99  # overfills the donor's Roche lobe: setup fails
        models.append(lc.Model.from_text(W.model_text(ov, base)))
    times = W.linspace_times(-0.2, 1.2, 400)
    obs = _obs(len(times), 9)
    dp = lc.DevicePackBatch(models, times, 0.001, 1, chunk=2048, limb=limb, **obs)
    for s in (0, 5, 17, 31, 47):
        d = dp.check(s)
        assert all(d[k] == 0 for k in CHECK_KEYS), (s, d)
        if limb and s != 17:
            assert d["host_limb_items"] > 0, (s, d)
    r1 = dp.run()
    assert np.isnan(r1["chisq"][17]) and np.all(np.isfinite(np.delete(r1["chisq"], 17)))
    _host_limb(lc, limb)
    lc.set_tuning(2048)
    rh = lc.Batch(models, times, 0.001, 1, **obs).run()
    ok = np.isfinite(rh["chisq"])
    assert not ok[17] and ok.sum() == 47
    assert np.array_equal(r1["chisq"][ok], rh["chisq"][ok])
    assert np.array_equal(r1["flux"][ok], rh["flux"][ok])
    dp.upload()
    r2 = dp.run()
    assert np.array_equal(r2["chisq"][ok], r1["chisq"][ok]) and np.array_equal(r2["flux"][ok], r1["flux"][ok])
    ms_pack, ms_flux, ms_total = dp.time(reps=3)
    assert 0.0 < ms_pack and 0.0 < ms_flux and ms_pack + ms_flux <= ms_total
    # Two work-table sets: consecutive launches alternate between them (the
    # pack of one beside the flux pass of the other); each gives the same.
    dp.set_slots(2)
    for _ in range(3):
        dp.launch()
    dp.sync()
    for _ in range(2):
        r5 = dp.run()
        assert np.array_equal(r5["chisq"][ok], r1["chisq"][ok]) and np.array_equal(r5["flux"][ok], r1["flux"][ok])
    d = dp.check(5)
    assert all(d[k] == 0 for k in CHECK_KEYS), d
    ms_pack2, ms_flux2, ms_total2 = dp.time(reps=3)
    assert 0.0 < ms_pack2 and 0.0 < ms_flux2 and 0.0 < ms_total2
    st = dp.stats()
    assert st["n_sets"] == 48 and st["n_valid"] == 47 and st["n_items"] > 0
    assert st["input_bytes"] > 0 and st["upload_s"] > 0.0
    hs = lc.Batch(models, times, 0.001, 1, **obs).stats()
    for k in ("pairs_evaluated", "pairs_by_moments", "moment_evals", "visits", "pairs_limb"):
        assert st[k] == hs[k], (k, st[k], hs[k])
    assert st["n_limb_chunks"] == hs["limb_chunks"] and st["n_limb_items"] == hs["limb_items"]
    assert (st["n_limb_chunks"] > 0) == (limb > 0)
    # new observations: the same as a batch built with them
    obs2 = _obs(len(times), 10)
    f2 = np.tile(obs2["flux"], (48, 1))
    dp.set_obs(flux2d=f2)
    r3 = dp.run()
    r4 = lc.Batch(models, times, 0.001, 1, flux2d=f2, ferr2d=np.tile(obs["ferr"], (48, 1)),
                  weight2d=np.tile(obs["weight"], (48, 1))).run()
    assert np.array_equal(r3["chisq"][ok], r4["chisq"][ok])


def test_device_pack_many_grids(host_tuning):
    """1600 sets, over 4096 segments: each grid is sorted by its own
    workgroup, the largest first, so grids finish out of order. Tables
    equal the host's on sets at both ends and in the middle, and every
    result equals the host-packed batch's."""
    lc = host_tuning
    base = W.chisq_base_text()
    n = 1600
    p = W.chisq_sets(n, seed=11, nlat_lo=12, nlat_hi=20)
    models = [lc.Model.from_text(W.model_text(
        {"q": s[0], "iangle": s[1], "r1": s[2], "t1": s[3], "t2": s[4], "nlat1f": 8, "nlat1c": 8,
         "nlat2f": int(s[5]), "nlat2c": int(s[5])}, base)) for s in p]
    times = W.linspace_times(-0.2, 1.2, 128)
    obs = _obs(len(times), 12)
    dp = lc.DevicePackBatch(models, times, 0.001, 1, chunk=1024, **obs)
    st = dp.stats()
    assert st["n_segs"] > 4096, st["n_segs"]
    for s in (0, 1, n // 2, n - 2, n - 1):
        d = dp.check(s)
        assert all(d[k] == 0 for k in CHECK_KEYS), (s, d)
    rd = dp.run()
    _host_limb(lc, 0)
    lc.set_tuning(1024)
    rh = lc.Batch(models, times, 0.001, 1, **obs).run()
    ok = np.isfinite(rh["chisq"])
    assert ok.sum() > n // 2
    assert np.array_equal(np.isfinite(rd["chisq"]), ok)
    assert np.array_equal(rd["chisq"][ok], rh["chisq"][ok])
    assert np.array_equal(rd["flux"][ok], rh["flux"][ok])


@pytest.mark.timeout(600)
def test_device_pack_chisq_4096_full_size(host_tuning):
    """The chi-squared workload at its full size: 4096 sets, 2000 times,
    star-2 grids of 10k-150k elements, chunk 4096. Tables equal the host's
    on sets at both ends and in the middle; a second launch gives the same
    values bit for bit; curves and autoscaled chi-squared on a sample of
    sets meet the parity bar against the FP64 oracle (1e-6 relative on the
    flux, 1e-5 on chi-squared)."""
    lc = host_tuning
    n = 4096
    base = W.chisq_base_text()
    texts = [W.model_text({"q": s[0], "iangle": s[1], "r1": s[2], "t1": s[3], "t2": s[4],
                           "nlat2f": int(s[5]), "nlat2c": int(s[5])}, base) for s in W.chisq_sets(n)]
    times = W.linspace_times(-0.2, 1.2, 2000)
    models = [lc.Model.from_text(x) for x in texts]
    dp = lc.DevicePackBatch(models, times, 0.001, 1, chunk=4096)
    st = dp.stats()
    assert st["n_sets"] == n and st["n_valid"] == n, st
    print(f"chisq_4096: {st['input_bytes'] / 1e9:.2f} GB of raw grids", flush=True)
    true = dp.run(scale=False)["flux"]
    assert np.all(np.isfinite(true)) and np.all(true > 0.0)
    rng = np.random.default_rng(4242)
    obs = true * (1.0 + 0.01 * rng.standard_normal(true.shape))
    sig = 0.01 * np.abs(true)
    dp.set_obs(obs, sig, np.ones_like(obs))
    for s in (0, n // 2, n - 1):
        d = dp.check(s)
        assert all(d[k] == 0 for k in CHECK_KEYS), (s, d)
    r1 = dp.run()
    r2 = dp.run()
    assert np.array_equal(r1["flux"], r2["flux"]) and np.array_equal(r1["chisq"], r2["chisq"])
    assert np.all(np.isfinite(r1["chisq"])) and np.all(r1["chisq"] > 0.0)
    idx = np.array([0, 1365, 2730, n - 1])
    ob = pyoracle.OracleBatch([texts[i] for i in idx], times, 0.001, 1, nthreads=16,
                              flux2d=obs[idx], ferr2d=sig[idx], weight2d=np.ones((len(idx), len(times))))
    chi, flux = ob.flux(scale=True, nthreads=16)
    assert np.max(np.abs(r1["flux"][idx] / flux - 1)) <= 1e-6
    assert np.max(np.abs(r1["chisq"][idx] / chi - 1)) <= 1e-5
