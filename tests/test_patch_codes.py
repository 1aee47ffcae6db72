"""CPU checks of the engine's exact work skipping.

The engine decides per (64-element patch, 64-lane tile) whether to skip a
patch (it adds exactly zero to every lane), evaluate it from its direction
moments (every element seen whole by every lane), drop the eclipse test (no
lane eclipsed) or evaluate it fully. Model.selfcheck_patch_codes re-evaluates
every (element, lane) pair in FP64 both as coded and by brute force; any wrong
decision shows up as a difference of the order of one element's term, far
above the 1e-10 bound used here. Dense times (2000 over 1.4 cycles) give
narrow tiles, so every code occurs.
"""
import numpy as np
import pytest

from lcurve_rs import workloads as W
from variants import VARIANTS

DENSE = W.linspace_times(-0.2, 1.2, 2000)
CODES = ("full", "no_eclipse", "skip", "moments")


def _pairs(counts):
    return sum(counts[k] for k in CODES)


@pytest.fixture(scope="module")
def lc():
    import lcurve_rs
    return lcurve_rs


def test_golden_model_codes(lc, golden_model_text):
    m = lc.Model.from_text(golden_model_text)
    rel, counts = m.selfcheck_patch_codes(DENSE)
    assert rel < 1e-10
    assert counts["skip"] > 0 and counts["moments"] > 0 and counts["no_eclipse"] > 0
    assert counts["full"] > 0  # the white dwarf's eclipse by the donor
    # No-eclipse pairs are split per lane: whole (moments), part, none.
    assert counts["split_whole"] > 0 and counts["split_part"] > 0 and counts["split_none"] > 0


@pytest.mark.parametrize("name", sorted(VARIANTS))
def test_variant_codes(lc, golden_model_text, name):
    m = lc.Model.from_text(W.model_text(VARIANTS[name], golden_model_text))
    rel, counts = m.selfcheck_patch_codes(DENSE)
    assert rel < 1e-10, (name, rel, counts)
    assert _pairs(counts) > 0


def test_sparse_times_skip_less(lc, golden_model_text):
    """Tiles spanning much of an orbit leave little to skip; the check still holds."""
    m = lc.Model.from_text(golden_model_text)
    rel, counts = m.selfcheck_patch_codes(W.linspace_times(-0.2, 1.2, 100))
    assert rel < 1e-10
    assert counts["skip"] + counts["moments"] < counts["no_eclipse"] + counts["full"]


def test_exposure_smearing_codes(lc, golden_model_text):
    m = lc.Model.from_text(golden_model_text)
    rel, _ = m.selfcheck_patch_codes(W.linspace_times(-0.2, 1.2, 600), expose=0.01, ndiv=5)
    assert rel < 1e-10


def test_chisq_workload_set_codes(lc):
    s = W.chisq_sets(1)[0]
    text = W.model_text({"q": s[0], "iangle": s[1], "r1": s[2], "t1": s[3], "t2": s[4],
                         "nlat2f": int(s[5]), "nlat2c": int(s[5])}, W.chisq_base_text())
    rel, counts = lc.Model.from_text(text).selfcheck_patch_codes(DENSE)
    assert rel < 1e-10
    tot = _pairs(counts)
    # most patch-tile pairs avoid the element-by-element loop
    assert (counts["skip"] + counts["moments"]) / tot > 0.5
    # and of the no-eclipse pairs' lanes, a minority straddles the limb
    split = counts["split_whole"] + counts["split_part"] + counts["split_none"]
    assert 0 < counts["split_part"] < 0.5 * split


def test_disc_spot_codes(lc):
    text, _, _, _ = W.disc_spot(0.3, 100)
    rel, counts = lc.Model.from_text(text).selfcheck_patch_codes(DENSE[::2], expose=0.005, ndiv=5)
    assert rel < 1e-10
    assert counts["moments"] > 0 and counts["split_whole"] > 0
