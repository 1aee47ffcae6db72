"""GPU test of lch_flux_eval_static, the entry point that allocates nothing:
through the C ABI, its curves equal lch_flux_eval's bit for bit on two calls
in a row (the second smaller, so leftovers of the first in the static
scratch must not leak into it), a call too large for the scratch fails
with LCH_ERR_INVALID and a clear message, and the next call still works."""
import ctypes

import numpy as np
import pytest

from lcurve_rs import _native
from lcurve_rs import workloads as W

pytestmark = pytest.mark.gpu


def _static(lc, gm, phase, expose, ndiv, slfac, third):
    n = len(phase)
    ph = np.ascontiguousarray(phase, dtype=np.float64)
    ex = np.ascontiguousarray(np.broadcast_to(expose, (n,)), dtype=np.float64)
    nd = np.ascontiguousarray(np.broadcast_to(ndiv, (n,)), dtype=np.int32)
    sl = np.ascontiguousarray(np.broadcast_to(slfac, (n,)), dtype=np.float64)
    out = np.full(n, np.nan)
    L = _native.lib()
    L.lch_flux_eval_static.restype = ctypes.c_int
    L.lch_flux_eval_static.argtypes = [ctypes.c_void_p, ctypes.c_void_p, _native.D, _native.D, _native.I,
                                       _native.D, ctypes.c_double, _native.D, ctypes.c_int]
    rc = L.lch_flux_eval_static(lc._ctx(), gm._h, ph.ctypes.data_as(_native.D), ex.ctypes.data_as(_native.D),
                                nd.ctypes.data_as(_native.I), sl.ctypes.data_as(_native.D), float(third),
                                out.ctypes.data_as(_native.D), n)
    return rc, out


def test_static_entry_point_matches_flux_eval(lcurve_rs, golden_model_text):
    from lcurve_rs import grids as G
    lc = lcurve_rs
    big = W.model_text({"nlat1f": 40, "nlat1c": 20, "nlat2f": 90, "nlat2c": 45}, golden_model_text)
    small = W.model_text({"nlat1f": 12, "nlat1c": 12, "nlat2f": 20, "nlat2c": 20}, golden_model_text)
    calls = [(big, W.linspace_times(-0.2, 1.2, 900), 0.003, 3), (small, W.linspace_times(0.3, 0.7, 70), 0.001, 1)]
    for text, times, expose, ndiv in calls:
        m = lc.Model.from_text(text)
        s = G.prepare(m, times, expose, ndiv)
        gm = G.upload_setup(s)
        rc, got = _static(lc, gm, s.phase, s.expose, s.ndiv, s.slfac, s.third)
        assert rc == 0, _native.last_error()
        ref = gm.flux(s.phase, s.expose, s.ndiv, s.slfac, s.third)
        assert np.array_equal(got, ref), np.nanmax(np.abs(got / ref - 1))
        # and lch_light_curve's curve (same kernels, same packing)
        assert np.array_equal(got, m.light_curve(times=times, expose=expose, ndiv=ndiv).calc)


def test_static_entry_point_rejects_oversized_calls(lcurve_rs, golden_model_text):
    from lcurve_rs import grids as G
    lc = lcurve_rs
    m = lc.Model.from_text(W.model_text({"nlat1f": 12, "nlat1c": 12, "nlat2f": 20, "nlat2c": 20}, golden_model_text))
    t = W.linspace_times(-0.2, 1.2, 60)
    s = G.prepare(m, t, 0.001, 1)
    gm = G.upload_setup(s)
    # a million sub-phases need ~100 MB of lane and time tables
    n = 1_000_000
    ph = np.linspace(-0.2, 1.2, n)
    rc, _ = _static(lc, gm, ph, 0.0, 1, 1.0, 0.0)
    assert rc == _native.LCH_ERR_INVALID
    msg = _native.last_error()
    assert "LCH_STATIC_SCRATCH_BYTES" in msg and "needs more than" in msg, msg
    rc, got = _static(lc, gm, s.phase, s.expose, s.ndiv, s.slfac, s.third)
    assert rc == 0 and np.array_equal(got, gm.flux(s.phase, s.expose, s.ndiv, s.slfac, s.third))
