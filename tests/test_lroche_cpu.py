"""CPU checks of the lroche front end (lcurve_rs.lroche, src/main.rs): the
reference prints floats with Rust's formatting, which differs from C's and
Python's defaults, and reports errors as `lroche: <message>` with exit
status 1. The expected strings below are Rust's documented output forms:
`{:.Ne}` writes the exponent without '+' or padding, `{:W.Ne}` right-aligns,
`{}` writes the shortest round-trip decimal without an exponent."""
import numpy as np
import pytest

from lcurve_rs import lroche


@pytest.mark.parametrize("x, prec, width, expect", [
    (0.001234, 12, 0, "1.234000000000e-3"),
    (-0.2, 15, 20, "-2.000000000000000e-1"),
    (1.0, 15, 20, " 1.000000000000000e0"),
    (7.488209078e-10, 10, 17, " 7.4882090780e-10"),
    (123456.0, 12, 0, "1.234560000000e5"),
    (0.0, 12, 0, "0.000000000000e0"),
    (-0.0, 3, 0, "-0.000e0"),
    (float("nan"), 12, 0, "NaN"),
    (float("inf"), 10, 17, "              inf"),
    (-float("inf"), 12, 0, "-inf"),
    (1e300, 2, 0, "1.00e300"),
])
def test_rust_exponent_format(x, prec, width, expect):
    assert lroche.fmt_exp(x, prec, width) == expect


@pytest.mark.parametrize("x, expect", [
    (0.001, "0.001"), (1.0, "1"), (0.0, "0"), (-0.0, "-0"), (1e-7, "0.0000001"),
    (0.1 + 0.2, "0.30000000000000004"), (1e21, "1000000000000000000000"), (-2.5, "-2.5"),
    (float("nan"), "NaN"), (float("inf"), "inf"),
])
def test_rust_display_format(x, expect):
    assert lroche.fmt_float(x) == expect


def test_shortest_round_trip():
    rng = np.random.default_rng(3)
    for x in rng.standard_normal(200) * 10.0 ** rng.integers(-8, 8, 200):
        s = lroche.fmt_float(x)
        assert "e" not in s and float(s) == x


def test_errors_print_lroche_prefix_and_exit_1(tmp_path, capsys):
    rc = lroche.main([str(tmp_path / "missing_model.dat"), "none"])
    err = capsys.readouterr().err
    assert rc == 1 and err.startswith("lroche: ")


def test_command_line_defaults():
    a = lroche._parser().parse_args(["m.dat", "none"])
    assert (a.time1, a.time2, a.ntime, a.expose, a.ndivide, a.output, a.scale) == (-0.2, 0.8, 500, 0.001, 1,
                                                                                 None, False)
    a = lroche._parser().parse_args(["m.dat", "d.dat", "--scale", "-o", "out.txt", "--ntime", "7"])
    assert a.scale and a.output == "out.txt" and a.ntime == 7
