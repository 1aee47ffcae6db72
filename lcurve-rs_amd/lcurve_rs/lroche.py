"""lroche: the light curve of a white dwarf / Roche-distorted star binary,
computed by the GPU engine, with the reference's command line and output
(src/main.rs:6-120):

    python -m lcurve_rs.lroche MODEL DATA [--time1 T1] [--time2 T2]
        [--ntime N] [--expose E] [--ndivide N] [-o OUTPUT] [--scale]

DATA is a data file (time expose flux ferr weight ndiv per line) or 'none'
(any case) for ``ntime`` regularly spaced times from ``time1`` to ``time2``
with exposure ``expose`` over ``ndivide`` sub-times, flux and ferr 0 and
weight 1. ``--scale`` autoscales (only with a data file). The summary
(chi-squared with a data file, white dwarf's contribution, log g and
volume-averaged radii) goes to stdout, and the curve, one line per time
(time expose calc ferr weight ndiv), to OUTPUT or stdout. Numbers are
printed as the reference prints them: ``{:.Ne}`` as Rust's exponent form
(no '+', no exponent padding), ``{}`` of a float as its shortest
round-trip decimal without an exponent. Errors print ``lroche: <message>``
to stderr and exit with status 1.
"""
import argparse
import math
import sys

import numpy as np


def fmt_exp(x, prec, width=0):
    """Rust's ``{:width.prec e}`` of an f64: d.ddd...e<exp>, right-aligned."""
    x = float(x)
    if math.isnan(x):
        s = "NaN"
    elif math.isinf(x):
        s = "inf" if x > 0 else "-inf"
    else:
        m, e = f"{x:.{prec}e}".split("e")
        s = f"{m}e{int(e)}"
    return s.rjust(width)


def fmt_float(x):
    """Rust's ``{}`` of an f64: the shortest decimal that reads back as x,
    written out without an exponent."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "inf" if x > 0 else "-inf"
    return np.format_float_positional(x, unique=True, trim="-")


def _parser():
    ap = argparse.ArgumentParser(prog="lroche",
                                 description="Compute light curve of a white dwarf / Roche-distorted star binary")
    ap.add_argument("model", help="Model parameter file")
    ap.add_argument("data", help="Data file (use 'none' to generate regularly-spaced times)")
    ap.add_argument("--time1", type=float, default=-0.2, help="First time to compute (when data = 'none')")
    ap.add_argument("--time2", type=float, default=0.8, help="Last time to compute (when data = 'none')")
    ap.add_argument("--ntime", type=int, default=500, help="Number of times to compute (when data = 'none')")
    ap.add_argument("--expose", type=float, default=0.001, help="Exposure time (when data = 'none')")
    ap.add_argument("--ndivide", type=int, default=1, help="Number of exposure subdivisions (when data = 'none')")
    ap.add_argument("-o", "--output", default=None, help="Output file for computed light curve")
    ap.add_argument("--scale", action="store_true", help="Autoscale to minimise chi-squared (only with real data)")
    return ap


def run(argv=None, out=None):
    """Run lroche with command-line arguments ``argv``; the printed text
    goes to ``out`` (default stdout). Returns the LcResult."""
    import lcurve_rs as lc
    from . import _native

    out = sys.stdout if out is None else out
    a = _parser().parse_args(argv)
    if not lc.has_hip():
        raise RuntimeError("no HIP device: the engine has no CPU path")
    lc.set_device("hip:0")
    mdl = lc.Model(a.model)
    no_file = a.data.lower() == "none"
    if no_file:
        if a.ntime <= 0:
            raise ValueError("No data points")
        # time1 + (time2 - time1) * i / (ntime - 1), in that order; one
        # time divides 0 by 0 there, as the reference does.
        with np.errstate(invalid="ignore"):
            t = np.array([a.time1 + np.float64((a.time2 - a.time1) * i) / np.float64(a.ntime - 1)
                          for i in range(a.ntime)])
        n = len(t)
        cols = {"time": t, "expose": np.full(n, a.expose), "flux": np.zeros(n), "ferr": np.zeros(n),
                "weight": np.ones(n), "ndiv": np.full(n, a.ndivide, dtype=np.int32)}
    else:
        cols = lc.read_data(a.data)
        if len(cols["time"]) == 0:
            raise ValueError("No data points")
    scale = a.scale and not no_file
    rdata = not no_file
    res = _light_curve(lc, _native, mdl, cols, scale, rdata)

    if rdata:
        print(f"Weighted chi**2 = {fmt_exp(res.chisq, 12)}, wnok = {fmt_exp(res.wnok, 12)}", file=out)
        if mdl.iscale:
            print("Scale factors = " + ", ".join(fmt_exp(v, 12) for v in res.sfac[:5]), file=out)
        else:
            print(f"Scale factor = {fmt_exp(res.sfac[0], 12)}", file=out)
    print(f"White dwarf's contribution = {fmt_exp(res.wdwarf, 12)}", file=out)
    print(f"log10(g1 [cgs]) = {fmt_exp(res.logg1, 12)}", file=out)
    print(f"log10(g2 [cgs]) = {fmt_exp(res.logg2, 12)}", file=out)
    print(f"Vol-averaged r1 = {fmt_exp(res.rv1, 12)}", file=out)
    print(f"Vol-averaged r2 = {fmt_exp(res.rv2, 12)}", file=out)

    lines = [f"{fmt_exp(cols['time'][i], 15, 20)} {fmt_float(cols['expose'][i])} {fmt_exp(res.calc[i], 10, 17)} "
             f"{fmt_float(cols['ferr'][i])} {fmt_float(cols['weight'][i])} {int(cols['ndiv'][i])}"
             for i in range(len(cols["time"]))]
    if a.output:
        with open(a.output, "w") as f:
            for line in lines:
                f.write(line + "\n")
        print(f"Written data to {a.output}", file=out)
    else:
        for line in lines:
            print(line, file=out)
    return res


def _light_curve(lc, _native, mdl, cols, scale, rdata):
    import ctypes
    n = len(cols["time"])
    calc = np.zeros(n)
    res = _native.LcResultC()
    d = {k: np.ascontiguousarray(cols[k], dtype=np.float64) for k in ("time", "expose", "flux", "ferr", "weight")}
    nd = np.ascontiguousarray(cols["ndiv"], dtype=np.int32)
    dp = lambda v: v.ctypes.data_as(_native.D)
    _native.check(_native.lib().lch_light_curve(
        lc._ctx(), mdl._h, dp(d["time"]), dp(d["expose"]), dp(d["flux"]), dp(d["ferr"]), dp(d["weight"]),
        nd.ctypes.data_as(_native.I), n, int(bool(scale)), int(rdata), dp(calc), ctypes.byref(res)))
    return lc.LcResult(d["time"].copy(), calc, res)


def main(argv=None):
    try:
        run(argv)
    except SystemExit:
        raise
    except Exception as e:  # the reference's `lroche: {e}` and exit status 1
        print(f"lroche: {e}", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
