"""The lroche front end on the GPU (src/main.rs): its curve is the
engine's light curve, within the golden tolerance of the reference's
500-point curve, printed in the reference's layout."""
import io
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

MODEL = os.path.join(GOLDEN, "test_model.dat")


def _curve_lines(text):
    rows = [ln.split() for ln in text.splitlines() if len(ln.split()) == 6]
    return rows


def test_lroche_generated_times(lcurve_rs, golden_lc):
    from lcurve_rs import lroche
    buf = io.StringIO()
    lroche.run([MODEL, "none", "--time2", "1.2"], out=buf)
    text = buf.getvalue()
    head = text.splitlines()[:5]
    assert head[0].startswith("White dwarf's contribution = ")
    assert [h.split(" = ")[0] for h in head[1:]] == ["log10(g1 [cgs])", "log10(g2 [cgs])", "Vol-averaged r1",
                                                    "Vol-averaged r2"]
    assert "Weighted chi**2" not in text
    rows = _curve_lines(text)
    assert len(rows) == 500
    t = np.array([float(r[0]) for r in rows])
    calc = np.array([float(r[2]) for r in rows])
    assert all(r[1] == "0.001" and r[3] == "0" and r[4] == "1" and r[5] == "1" for r in rows)
    assert np.allclose(t, golden_lc[:, 0], rtol=0, atol=1e-15)
    res = lcurve_rs.Model(MODEL).light_curve(time1=-0.2, time2=1.2, ntime=500, expose=0.001, ndiv=1)
    # The curve is printed with 11 significant digits.
    assert np.allclose(calc, res.calc, rtol=6e-11, atol=0)
    assert np.max(np.abs(calc / golden_lc[:, 2] - 1.0)) < 1e-6


def test_lroche_with_data_file_and_output(lcurve_rs, tmp_path):
    from lcurve_rs import lroche
    data = tmp_path / "obs.dat"
    rng = np.random.default_rng(2)
    t = np.linspace(-0.2, 0.8, 200)
    res0 = lcurve_rs.Model(MODEL).light_curve(times=t, expose=0.002, ndiv=3)
    y = 3.0 * res0.calc * (1.0 + 0.01 * rng.standard_normal(len(t)))
    ferr = 0.01 * 3.0 * res0.calc
    ferr[5] = -ferr[5]  # a negative error: weight 0, |ferr| printed
    with open(data, "w") as f:
        f.write("# time expose flux ferr weight ndiv\n")
        for i in range(len(t)):
            f.write(f"{float(t[i])!r} 0.002 {float(y[i])!r} {float(ferr[i])!r} 1 3\n")
    out = tmp_path / "curve.txt"
    buf = io.StringIO()
    res = lroche.run([MODEL, str(data), "--scale", "-o", str(out)], out=buf)
    lines = buf.getvalue().splitlines()
    assert lines[0].startswith("Weighted chi**2 = ") and ", wnok = " in lines[0]
    assert lines[1].startswith("Scale factor = ")
    assert lines[-1] == f"Written data to {out}"
    sfac = float(lines[1].split(" = ")[1])
    assert sfac == pytest.approx(3.0, rel=0.01)
    chi = float(lines[0].split(" = ")[1].split(",")[0])
    assert chi == pytest.approx(res.chisq, rel=1e-11) and 100 < chi < 400
    rows = _curve_lines(out.read_text())
    assert len(rows) == 200
    assert rows[5][4] == "0" and float(rows[5][3]) == pytest.approx(-ferr[5], rel=1e-15)
    assert all(r[5] == "3" and r[1] == "0.002" for r in rows)
    calc = np.array([float(r[2]) for r in rows])
    assert np.allclose(calc, res.calc, rtol=6e-11, atol=0)


def test_lroche_iscale_prints_five_factors(lcurve_rs, golden_model_text, tmp_path):
    """An iscale model (per-component fit, orchestration.rs:181-227) prints
    five scale factors, as src/main.rs does; they equal the LcResult's."""
    from lcurve_rs import lroche
    from lcurve_rs import workloads as W
    model = tmp_path / "iscale_model.dat"
    model.write_text(W.model_text({"iscale": 1}, golden_model_text))
    assert lcurve_rs.Model(str(model)).iscale and not lcurve_rs.Model(MODEL).iscale
    t = np.linspace(-0.2, 0.8, 150)
    base = lcurve_rs.Model(MODEL).light_curve(times=t, expose=0.001, ndiv=1)
    data = tmp_path / "obs.dat"
    with open(data, "w") as f:
        for i in range(len(t)):
            f.write(f"{float(t[i])!r} 0.001 {float(2.0 * base.calc[i])!r} {float(0.01 * base.calc[i])!r} 1 1\n")
    buf = io.StringIO()
    res = lroche.run([str(model), str(data), "--scale"], out=buf)
    line = buf.getvalue().splitlines()[1]
    assert line.startswith("Scale factors = ")
    vals = [float(v) for v in line.split(" = ")[1].split(", ")]
    assert len(vals) == 5
    assert np.allclose(vals, res.sfac[:5], rtol=1e-12, atol=0)
