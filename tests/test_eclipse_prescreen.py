"""CPU checks of the certified short cuts in the CPU setup.

ingress_egress (crates/roche/src/eclipse.rs:270-373) scans up to 1000 phases
per element with fblink until the line of sight first dips into the
eclipsing star's Roche lobe; the setup skips a scan step only when a rigorous
bound proves fblink false there. face() (surface.rs:54-121) chops for the
surface radius; the setup decides a midpoint's comparison without the
potential only when a certified Taylor bound fixes it. The grids must come
out bit-identical with and without both. Model.selfcheck_eclipse_prescreen
builds both stars' fine and coarse grids both ways and counts elements whose
position, normal, area, gravity or eclipse intervals differ in any bit.
"""
import numpy as np
import pytest

from lcurve_rs import workloads as W
from variants import VARIANTS


@pytest.fixture(scope="module")
def lc():
    import lcurve_rs
    return lcurve_rs


SMALL = {"nlat1f": 16, "nlat1c": 8, "nlat2f": 24, "nlat2c": 12}


def _check(lc, text, min_skipped=1):
    r = lc.Model.from_text(text).selfcheck_eclipse_prescreen()
    assert r["elements"] > 0
    assert r["mismatches"] == 0, r
    assert r["skipped"] >= min_skipped, r
    return r


def test_golden_model(lc, golden_model_text):
    r = _check(lc, golden_model_text, min_skipped=10000)
    # Most scan steps pass beside the donor's lobe and are certified.
    assert r["skipped"] > 5 * r["evaluated"], r


@pytest.mark.parametrize("name", ["roche_primary", "beaming_spin", "npole_infill"])
def test_variants(lc, golden_model_text, name):
    # roche_primary: star 2's elements eclipsed by a Roche-distorted star 1
    # (the Primary form of the bound); beaming_spin: asynchronous rotation.
    _check(lc, W.model_text(VARIANTS[name], golden_model_text))


def test_no_eclipses_untouched(lc, golden_model_text):
    r = lc.Model.from_text(W.model_text(VARIANTS["no_eclipses"], golden_model_text)) \
        .selfcheck_eclipse_prescreen()
    assert r["mismatches"] == 0 and r["skipped"] == 0 and r["evaluated"] == 0, r


def _random_sets(n, seed):
    rng = np.random.default_rng(seed)
    for _ in range(n):
        ov = dict(SMALL)
        ov["q"] = float(rng.uniform(0.05, 2.0))
        ov["iangle"] = float(rng.uniform(65.0, 90.0))
        ov["r1"] = float(rng.uniform(0.005, 0.06))
        ov["spin1"] = float(rng.uniform(0.5, 3.0))
        ov["spin2"] = float(rng.uniform(0.6, 1.4))
        ov["delta_phase"] = float(10.0 ** rng.uniform(-8, -5))
        # Lobe-filling donor, or an underfilling one (a fraction of the
        # Eggleton lobe radius keeps it inside its lobe).
        if rng.random() < 0.7:
            ov["r2"] = -1
        else:
            q = ov["q"]
            rl = 0.49 * q ** (2 / 3) / (0.6 * q ** (2 / 3) + np.log(1 + q ** (1 / 3)))
            ov["r2"] = float(rl * rng.uniform(0.6, 0.9))
        if rng.random() < 0.3:
            # Roche-distorted star 1 eclipsing star 2.
            ov["roche1"] = 1
            q = 1 / ov["q"]
            rl1 = 0.49 * q ** (2 / 3) / (0.6 * q ** (2 / 3) + np.log(1 + q ** (1 / 3)))
            ov["r1"] = float(rl1 * rng.uniform(0.5, 0.9))
        yield ov


@pytest.mark.parametrize("seed", range(4))
def test_random_parameter_sets(lc, golden_model_text, seed):
    n = 0
    for ov in _random_sets(6, seed):
        text = W.model_text(ov, golden_model_text)
        try:
            r = lc.Model.from_text(text).selfcheck_eclipse_prescreen()
        except lc.LcurveError:
            continue  # unphysical draw (a star larger than its lobe)
        assert r["mismatches"] == 0, (ov, r)
        n += 1
    assert n >= 3
