"""CPU tests of the C ABI: the library loads, exports every declared entry
point, and its host-only calls (model and data files) behave like the
reference's parser (crates/lcurve/src/model.rs, types.rs:229-254).
No compute call is made here."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

import lcurve_rs as lc
from lcurve_rs import _native, workloads as W

HEADER = os.path.join(ROOT, "include", "lcurve_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lch_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _native.lib()
    names = declared_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert set(declared_symbols()) == set(_native.SIGNATURES)


def test_library_reports_version_and_device_count():
    L = _native.lib()
    assert L.lch_version() >= 1
    assert L.lch_device_count() >= 0


def test_model_load_and_params():
    m = lc.Model(os.path.join(GOLDEN, "test_model.dat"))
    # model.rs tests: get / set / round trip, metadata kept
    assert m.get_param("q") == 0.5
    assert m.get_param("iangle") == 82.0
    m.set_param("q", 0.42)
    assert m.q == 0.42
    p = m.get_pparam("q")
    assert p["range"] == 0.1 and p["dstep"] == 0.01 and p["vary"] is False and p["defined"] is True
    with pytest.raises(ValueError):
        m.get_param("nonexistent")
    with pytest.raises(ValueError):
        m.set_param("nonexistent", 1.0)
    c = m.copy()
    c.set_param("q", 0.3)
    assert m.q == 0.42 and c.q == 0.3


def test_optional_params_default_and_are_undefined():
    m = lc.Model(os.path.join(GOLDEN, "test_model.dat"))
    # pdot / temp_edge / absorb_edge are absent from the file
    for name in ("pdot", "temp_edge", "absorb_edge"):
        p = m.get_pparam(name)
        assert p["value"] == 0.0 and p["defined"] is False and p["range"] == 1e-10


def test_free_params():
    txt = W.model_text({})
    txt = txt.replace("q                    = 0.5 0.1 0.01 0", "q                    = 0.5 0.1 0.01 1")
    m = lc.Model.from_text(txt)
    free = m.get_free_params()
    assert [f["name"] for f in free] == ["q"] and free[0]["range"] == 0.1
    m.set_pparam("iangle", vary=True, range=3.0)
    assert [f["name"] for f in m.get_free_params()] == ["q", "iangle"]


@pytest.mark.parametrize("edit,msg", [
    (lambda t: t + "bogus = 1\n", "was not recognised"),
    (lambda t: "\n".join(l for l in t.splitlines() if not l.startswith("iangle")), "was not defined"),
    (lambda t: t + "stsp11_long = 10 1 1 0\n", "stsp11"),
    (lambda t: t.replace("limb1                = Poly", "limb1                = Linear"), "bad limb1"),
    (lambda t: t.replace("glens1               = 0", "glens1               = 1").replace(
        "roche1               = 0", "roche1               = 1"), "glens1=1 and roche1=1"),
    (lambda t: t.replace("q                    = 0.5 0.1 0.01 0", "q = 0.5 0.1"), "too few fields"),
])
def test_model_parse_errors(edit, msg):
    with pytest.raises(ValueError, match=msg):
        lc.Model.from_text(edit(W.model_text({})))


def test_comments_and_escaped_hash():
    txt = W.model_text({}).replace("q                    = 0.5 0.1 0.01 0",
                                   "q                    = 0.25 0.1 0.01 0   # comment")
    assert lc.Model.from_text("# header\n" + txt).q == 0.25


def test_read_data_file(tmp_path):
    f = tmp_path / "d.dat"
    f.write_text("# t e f fe w n\n\n0.1 0.01 1.5 0.1 1 3\n0.2 0.01 1.6 -0.2 1 1\nshort line 1\n"
                 "0.3 0.02 1.7 0.3 0.5 2\n")
    d = lc.read_data(str(f))
    assert np.allclose(d["time"], [0.1, 0.2, 0.3])
    assert np.allclose(d["ferr"], [0.1, 0.2, 0.3])       # negative ferr -> |ferr|
    assert np.allclose(d["weight"], [1.0, 0.0, 0.5])     # ... and weight 0
    assert list(d["ndiv"]) == [3, 1, 2]


def test_whitedwarf_model_parses():
    m = lc.Model.whitedwarf(q=0.3, iangle=85.0)
    assert m.q == 0.3 and m.iangle == 85.0


def test_device_strings():
    with pytest.raises(ValueError):
        lc.set_device("cpu")
    with pytest.raises(ValueError):
        lc.set_device("hip:-1")
    with pytest.raises(ValueError):
        lc.set_device("tpu")
    lc.set_device("cuda:0")
    assert lc.get_device() == "hip:0"
    lc.set_device("hip")
    assert lc.get_device() == "hip:0"


def test_compute_without_device_fails_loudly():
    if lc.has_hip():
        pytest.skip("a HIP device is present")
    m = lc.Model(os.path.join(GOLDEN, "test_model.dat"))
    with pytest.raises(_native.LcurveError):
        m.light_curve(ntime=10)


def test_surrogate_likelihood_feature_count():
    """The GPU surrogate likelihood's polynomial has 1 + every monomial of
    degree 1..d in n variables: C(n + d, d) columns, as
    lcurve_rs.mcmc.poly_features builds them."""
    from math import comb
    from lcurve_rs import mcmc
    from lcurve_rs.surrogate import n_poly_features
    for nd in (1, 3, 5):
        for deg in range(5):
            assert n_poly_features(nd, deg) == comb(nd + deg, deg)
            assert n_poly_features(nd, deg) == mcmc.poly_features(np.zeros((1, nd)), deg).shape[1]
    with pytest.raises(ValueError):
        n_poly_features(0, 2)
    with pytest.raises(ValueError):
        n_poly_features(5, 7)


def test_setup_export_matches_the_cpu_setup(golden_model_text):
    """lch_model_prepare (no device): the exported grids are the CPU setup's
    (the oracle runs the same setup code), eclipse intervals in CSR form,
    coarse grids present only where the reference keeps separate ones, the
    phases those of orchestration.rs:231-237 and the flux parameters the
    model's."""
    import pyoracle
    from lcurve_rs import grids as G
    text = W.model_text({"deltat": 0.002, "pdot": 1e-5, "slope": 0.1, "third": 1e-12}, golden_model_text)
    t = W.linspace_times(-0.2, 1.2, 40)
    s = G.prepare(lc.Model.from_text(text), t, 0.002, 3)
    assert set(s.grids) == {"star1f", "star1c", "star2f", "star2c"}
    ob = pyoracle.OracleBatch(text, t, 0.002, 3)
    for name, gi in (("star1f", 0), ("star1c", 1), ("star2f", 2), ("star2c", 3)):
        o = ob.grid(0, gi)
        g = s.grids[name]
        assert np.array_equal(g["posn"], o["posn"]) and np.array_equal(g["dirn"], o["dirn"])
        assert np.array_equal(g["flux"].astype(np.float64), o["flux"])
        assert np.array_equal(np.diff(g["ecl_off"]), o["necl"])
        assert g["ecl_off"][0] == 0 and len(g["ingress"]) == g["ecl_off"][-1] == len(g["egress"])
        assert np.all(g["egress"] > g["ingress"])
    assert np.array_equal(s.phase, pyoracle.phases(text, t))
    assert np.all(s.expose == 0.002) and np.all(s.ndiv == 3) and s.third == 1e-12
    mid, rng = (t.min() + t.max()) / 2, (t.max() - t.min()) / 2
    assert np.allclose(s.slfac, 1 + 0.1 * (t - mid) / rng, rtol=1e-15)
    p = s.params
    assert p["q"] == 0.5 and p["iangle"] == 82.0 and p["gint_mode"] == 1
    assert p["limb1"] == 0 and p["ldc1"][:2] == [0.4, 0.0] and p["ldc2"][0] == 0.6
    # a model whose coarse grids are its fine ones exports no coarse grid
    s2 = G.prepare(lc.Model.from_text(W.model_text({"nlat1c": 20, "nlat2c": 40}, golden_model_text)), t)
    assert set(s2.grids) == {"star1f", "star2f"}
    # iscale models are fitted per component: not this boundary
    with pytest.raises(ValueError, match="iscale"):
        G.prepare(lc.Model.from_text(W.model_text({"iscale": 1}, golden_model_text)), t)
    assert "iscale" in _native.last_error()
