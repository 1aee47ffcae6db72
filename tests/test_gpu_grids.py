"""GPU tests of the grid-level boundary (lch_model_upload / lch_flux_eval).

* Grids exported from the engine's own CPU setup (lch_model_prepare) and
  uploaded as arrays give exactly lch_light_curve's curve, bit for bit:
  the same records, patches, work codes and kernels, with the time lanes'
  tiles independent of the calibration lanes.
* Hand-made grids (a few elements with chosen eclipse intervals: wrap-around
  past phase 1, the always-eclipsed (0, 1.1) pair, interval edges) against a
  brute-force comp_light written from flux.rs:7-111 and types.rs:48-125, at
  the FP32 element-term tolerance 1e-6.
* Malformed grids are rejected with LCH_ERR_INVALID and a per-context
  message (lch_last_error(ctx)).
"""
import numpy as np
import pytest

from lcurve_rs import workloads as W
from variants import VARIANTS

pytestmark = pytest.mark.gpu


def grids_mod():
    from lcurve_rs import grids
    return grids


BIT_EXACT_CASES = {
    "golden (fine/coarse Ginterp)": ({}, 0.001, 1),
    "no Ginterp": ({"nlat1c": 20, "nlat2c": 40}, 0.001, 1),
    "ndiv 5": ({}, 0.01, 5),
    "claret": (VARIANTS["claret_ld"], 0.001, 1),
    "beaming": (VARIANTS["beaming_spin"], 0.001, 1),
    "lensing": (VARIANTS["lensing"], 0.001, 1),
    "disc edge spot": (VARIANTS["disc_edge_spot"], 0.002, 3),
    "opaque disc, calibration copies": (VARIANTS["opaque_disc_calibration"], 0.001, 1),
    "ephemeris and polynomial factor": (dict(VARIANTS["ephemeris"], **VARIANTS["poly_factor"]), 0.001, 1),
}


@pytest.mark.parametrize("case", list(BIT_EXACT_CASES))
def test_uploaded_setup_matches_light_curve_bit_for_bit(lcurve_rs, golden_model_text, case):
    ov, expose, ndiv = BIT_EXACT_CASES[case]
    text = W.model_text(ov, golden_model_text)
    times = W.linspace_times(-0.2, 1.2, 333)
    m = lcurve_rs.Model.from_text(text)
    ref = m.light_curve(times=times, expose=expose, ndiv=ndiv).calc
    G = grids_mod()
    s = G.prepare(m, times, expose, ndiv)
    gm = G.upload_setup(s)
    got = gm.flux(s.phase, s.expose, s.ndiv, s.slfac, s.third)
    assert np.array_equal(got, ref), np.max(np.abs(got / ref - 1))
    # Calls at the same phases reuse the work tables with new polynomial
    # factors and third light: slfac * x + third exactly, x the bare sum.
    x = gm.flux(s.phase, s.expose, s.ndiv, np.ones(len(times)), 0.0)
    sl2, th2 = 0.5 + 0.01 * np.arange(len(times)), 1e-3 * float(np.abs(ref).max())
    assert np.array_equal(gm.flux(s.phase, s.expose, s.ndiv, sl2, th2), sl2 * x + th2)
    # and new phases rebuild them
    sh = gm.flux(s.phase + 0.25, s.expose, s.ndiv, s.slfac, s.third)
    assert not np.array_equal(sh, got) and np.array_equal(gm.flux(s.phase, s.expose, s.ndiv, s.slfac, s.third), got)


def test_given_ginterp_factors_equal_calibrated_ones(lcurve_rs, golden_model_text):
    """gint_mode 0: the caller's Ginterp factors are applied as given. Unit
    factors leave the fine / coarse switch uncalibrated (a small step at
    phase1 against the calibrated curve); s11 = s12 = 2 doubles star 1's
    coarse-grid light between phase1 and 1 - phase1 and changes nothing
    outside."""
    G = grids_mod()
    times = W.linspace_times(-0.2, 1.2, 200)
    m = lcurve_rs.Model.from_text(golden_model_text)
    s = G.prepare(m, times, 0.001, 1)
    cal = G.upload_setup(s).flux(s.phase, s.expose, s.ndiv, s.slfac, s.third)
    p = dict(s.params, gint_mode=0, scale11=1.0, scale12=1.0, scale21=1.0, scale22=1.0)
    ones = G.GridModel(s.grids, p).flux(s.phase, s.expose, s.ndiv, s.slfac, s.third)
    # with unit factors the fine/coarse switch shows as a jump at phase1
    rel = np.abs(ones / cal - 1)
    assert rel.max() > 1e-6
    assert rel.max() < 5e-2
    # scaling a coarse-grid stretch by s11 = 2 doubles star 1's share there
    p2 = dict(p, scale11=2.0, scale12=2.0)
    two = G.GridModel(s.grids, p2).flux(s.phase, s.expose, s.ndiv, s.slfac, s.third)
    ph = s.phase - np.floor(s.phase)
    ramp = (ph > 0.05) & (ph < 0.95)
    assert np.all(two[ramp] > ones[ramp]) and np.array_equal(two[~ramp], ones[~ramp])


# ----------------------------------------------------- brute-force reference
def visible(phase, ecl):
    phi = phase - np.floor(phase)
    for ing, eg in ecl:
        if (phi >= ing and phi <= eg) or phi <= eg - 1.0:
            return False
    return True


def imu(mu, law, a):
    if mu <= 0.0:
        return 0.0
    mu = min(mu, 1.0)
    om = 1.0 - mu
    if law == 0:
        return 1.0 - om * (a[0] + om * (a[1] + om * (a[2] + om * a[3])))
    s = np.sqrt(mu)
    return 1.0 - (a[0] + a[1] + a[2] + a[3]) + s * (a[0] + s * (a[1] + s * (a[2] + s * a[3])))


def brute_comp_light(grids, p, phase, expose, ndiv):
    """comp_light (flux.rs:7-111) without beaming or lensing, Ginterp with
    the given factors (types.rs:129-202), pure Python."""
    ci = np.cos(2 * np.pi * p["iangle"] / 360.0)
    si = np.sin(2 * np.pi * p["iangle"] / 360.0)
    p1, p2 = p["phase1"], p["phase2"]

    def gtype(ph):
        pn = ph - np.floor(ph)
        if pn <= p1 or pn >= 1 - p1:
            return 1
        if (p1 < pn < p2) or (1 - p2 < pn < 1 - p1):
            return 2
        return 3

    def s1(ph):
        pn = ph - np.floor(ph)
        if pn <= p1 or pn >= 1 - p1:
            return 1.0
        return (p["scale11"] * (1 - p1 - pn) + p["scale12"] * (pn - p1)) / (1 - 2 * p1)

    def s2(ph):
        pn = ph - np.floor(ph)
        if p2 <= pn <= 1 - p2:
            return 1.0
        if pn < 0.5:
            return (p["scale22"] * (p2 - pn) + p["scale21"] * (pn + p2)) / (2 * p2)
        return (p["scale21"] * (1 + p2 - pn) + p["scale22"] * (pn - 1 + p2)) / (2 * p2)

    def elems(name):
        g = grids.get(name)
        if g is None:
            return []
        off = g.get("ecl_off")
        out = []
        for i in range(len(g["flux"])):
            ecl = [] if off is None else list(zip(g["ingress"][off[i]:off[i + 1]], g["egress"][off[i]:off[i + 1]]))
            out.append((np.asarray(g["dirn"][i]), float(g["flux"][i]), ecl))
        return out

    total = 0.0
    for nd in range(ndiv):
        if ndiv == 1:
            phi, w = phase, 1.0
        else:
            phi = phase + expose * (nd - (ndiv - 1) / 2.0) / (ndiv - 1)
            w = 0.5 if nd in (0, ndiv - 1) else 1.0
        e = np.array([si * np.cos(2 * np.pi * phi), -si * np.sin(2 * np.pi * phi), ci])
        t = gtype(phi)
        st1 = elems("star1f" if t == 1 or "star1c" not in grids else "star1c")
        st2 = elems("star2f" if t == 3 or "star2c" not in grids else "star2c")
        ss = 0.0
        for d, f, ecl in st1:
            mu = float(e @ d)
            if visible(phi, ecl) and mu > p["mucrit1"]:
                ss += mu * f * imu(mu, p["limb1"], p["ldc1"])
        ss *= s1(phi)
        ss2 = 0.0
        for d, f, ecl in st2:
            mu = float(e @ d)
            if visible(phi, ecl) and mu > p["mucrit2"]:
                ss2 += mu * f * imu(mu, p["limb2"], p["ldc2"])
        ss += s2(phi) * ss2
        for name in ("disc", "edge"):
            for d, f, ecl in elems(name):
                mu = float(e @ d)
                if mu > 0 and visible(phi, ecl):
                    om = 1 - mu
                    ss += mu * f * (1 - om * (p["lin_limb_disc"] + p["quad_limb_disc"] * om))
        for d, f, ecl in elems("spot"):
            mu = float(e @ d)
            if mu > 0 and visible(phi, ecl):
                ss += mu * f
        total += w * ss
    return total / max(1, ndiv - 1)


def hand_grid(rng, n, center, radius, ecl_fn, flux_scale=1.0):
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    posn = np.asarray(center) + radius * d
    flux = (flux_scale * rng.uniform(0.5, 1.5, n)).astype(np.float32)
    off, ing, egr = [0], [], []
    for i in range(n):
        for a, b in ecl_fn(i):
            ing.append(a)
            egr.append(b)
        off.append(len(ing))
    return {"posn": posn, "dirn": d, "flux": flux, "ecl_off": np.array(off, dtype=np.int64),
            "ingress": np.array(ing), "egress": np.array(egr)}


def hand_params(**kw):
    p = {"iangle": 80.0, "q": 0.4, "limb1": 0, "limb2": 1, "ldc1": [0.4, 0.1, 0.0, 0.0],
         "ldc2": [0.55, -0.35, 0.65, -0.25], "mucrit1": 0.0, "mucrit2": 0.0, "lin_limb_disc": 0.3,
         "quad_limb_disc": 0.1, "beam_factor1": 0.0, "beam_factor2": 0.0, "spin1": 1.0, "spin2": 1.0,
         "velocity_scale": 0.0, "glens1": 0, "rlens1": 0.0, "phase1": 0.05, "phase2": 0.45,
         "gint_mode": 0, "scale11": 1.02, "scale12": 0.98, "scale21": 1.01, "scale22": 0.99}
    p.update(kw)
    return p


def test_hand_made_grids_match_brute_force(lcurve_rs):
    """Every eclipse-interval case of Point::visible: intervals inside
    [0, 1), wrapping past 1 (egress > 1 eclipses [0, egress - 1] too), the
    always-eclipsed sentinel (0, 1.1), several intervals per element; fine
    and coarse grids with given Ginterp factors; disc, edge and spot; ndiv 1
    and 7."""
    rng = np.random.default_rng(11)
    cases = [lambda i: [], lambda i: [(0.1 + 0.01 * (i % 7), 0.2 + 0.01 * (i % 5))],
             lambda i: [(0.95, 1.05)], lambda i: [(0.0, 1.1)] if i % 9 == 0 else [],
             lambda i: [(0.3, 0.35), (0.6, 0.7), (0.9 + 0.001 * i % 0.05, 1.02)]]
    grids = {"star1f": hand_grid(rng, 300, (0, 0, 0), 0.02, lambda i: cases[i % 5](i)),
             "star1c": hand_grid(rng, 150, (0, 0, 0), 0.02, lambda i: cases[(i + 1) % 5](i)),
             "star2f": hand_grid(rng, 500, (1, 0, 0), 0.3, lambda i: cases[(i + 2) % 5](i)),
             "star2c": hand_grid(rng, 200, (1, 0, 0), 0.3, lambda i: cases[(i + 3) % 5](i)),
             "disc": hand_grid(rng, 120, (0, 0, 0), 0.2, lambda i: cases[(i + 4) % 5](i), 0.3),
             "edge": hand_grid(rng, 70, (0, 0, 0), 0.3, lambda i: cases[i % 3](i), 0.2),
             "spot": hand_grid(rng, 40, (0.2, 0.2, 0), 0.01, lambda i: cases[(i + 1) % 3](i), 2.0)}
    p = hand_params()
    G = grids_mod()
    gm = G.GridModel(grids, p)
    phase = np.concatenate([np.linspace(-0.3, 1.4, 120), [0.1, 0.2, 0.3, 0.35, 0.95, 1.05, 0.05, 0.45]])
    for ndiv, expose in ((1, 0.0), (7, 0.02)):
        got = gm.flux(phase, expose, ndiv)
        ref = np.array([brute_comp_light(grids, p, ph, expose, ndiv) for ph in phase])
        scale = np.max(np.abs(ref))
        assert np.max(np.abs(got - ref)) <= 1e-6 * scale, (ndiv, np.max(np.abs(got - ref)) / scale)


def test_malformed_grids_are_rejected(lcurve_rs):
    from lcurve_rs import _native
    G = grids_mod()
    rng = np.random.default_rng(2)
    good = hand_grid(rng, 10, (0, 0, 0), 0.02, lambda i: [(0.1, 0.2)])
    off = good["ecl_off"].copy()
    off[2], off[3] = off[3], off[2]   # decreases once, still ends at len(ingress)
    bad_off = dict(good, ecl_off=off)
    five = hand_grid(rng, 4, (1, 0, 0), 0.3, lambda i: [(0.1 * k, 0.1 * k + 0.05) for k in range(5)])
    for grids, msg in (({"star1f": good, "star2f": bad_off}, "must not decrease"),
                       ({"star1f": good, "star2f": five}, "at most 4")):
        with pytest.raises(ValueError, match=msg):
            G.GridModel(grids, hand_params())
        # the context records the message too
        from lcurve_rs import _ctx
        assert msg in _native.last_error(_ctx())
    with pytest.raises(ValueError, match="gint_mode"):
        G.GridModel({"star1f": good}, hand_params(gint_mode=3))
    gm = G.GridModel({"star1f": good}, hand_params())
    with pytest.raises(ValueError):
        gm.flux(np.zeros(0), 0.0, 1)
