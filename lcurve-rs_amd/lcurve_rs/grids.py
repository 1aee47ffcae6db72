"""Grid-level boundary: surface-element grids as arrays, resident on the GPU.

The reference's GPU backend flattens lcurve's ``Point`` grids on the host
(crates/lcurve-cuda/src/transfer.rs:134-159) and evaluates the flux sum on
them (crates/lcurve-cuda/kernels/flux_kernel.cu:534-555). Here:

  * ``prepare(model, times, expose, ndiv)`` runs the engine's CPU setup (the
    reference's method) and returns the grids as numpy arrays with the flux
    parameters and per-time terms (lch_model_prepare);
  * ``GridModel(grids, params)`` uploads grids of that form, from any source
    (lch_model_upload), and ``GridModel.flux(phase, expose, ndiv, slfac,
    third)`` returns slfac * comp_light + third at those phases
    (lch_flux_eval).

A grid is a dict with ``posn`` [n, 3], ``dirn`` [n, 3], ``flux`` [n]
(float32), optional ``area`` / ``gravity`` (float32) and the eclipse
intervals in CSR form: ``ecl_off`` [n + 1] (int64), ``ingress``, ``egress``.
"""
import ctypes

import numpy as np

from . import _native
from ._native import GRID_NAMES, LCH_NGRID, FluxParamsC, GridSoaC, check

PARAM_KEYS = [f[0] for f in FluxParamsC._fields_]


def _params_to_c(p):
    c = FluxParamsC()
    for k in PARAM_KEYS:
        v = p[k]
        if k in ("ldc1", "ldc2"):
            for i in range(4):
                getattr(c, k)[i] = float(v[i])
        elif k in ("limb1", "limb2", "glens1", "gint_mode"):
            setattr(c, k, int(v))
        else:
            setattr(c, k, float(v))
    return c


def _params_from_c(c):
    out = {}
    for k in PARAM_KEYS:
        v = getattr(c, k)
        out[k] = list(v) if k in ("ldc1", "ldc2") else v
    return out


class Setup:
    """CPU setup of one model (lch_model_prepare): ``grids`` (dict name ->
    grid dict, empty grids omitted), ``params`` (dict of lch_flux_params) and
    the per-time ``phase``, ``expose``, ``ndiv``, ``slfac`` and ``third``."""

    def __init__(self, model, times, expose=0.001, ndiv=1):
        t = np.ascontiguousarray(times, dtype=np.float64)
        n = len(t)
        ex = np.ascontiguousarray(np.broadcast_to(expose, (n,)), dtype=np.float64)
        nd = np.ascontiguousarray(np.broadcast_to(ndiv, (n,)), dtype=np.int32)
        L = _native.lib()
        h = ctypes.c_void_p()
        check(L.lch_model_prepare(model._h, t.ctypes.data_as(_native.D), ex.ctypes.data_as(_native.D),
                                  nd.ctypes.data_as(_native.I), n, ctypes.byref(h)))
        try:
            self.grids = {}
            for g, name in enumerate(GRID_NAMES):
                v = GridSoaC()
                check(L.lch_setup_grid(h, g, ctypes.byref(v)))
                if v.n == 0:
                    continue
                k = int(v.n)
                ne = int(v.ecl_off[k])
                arr = lambda p, cnt, dt: np.ctypeslib.as_array(p, shape=(cnt,)).astype(dt, copy=True)
                self.grids[name] = {
                    "posn": arr(v.posn, 3 * k, np.float64).reshape(k, 3),
                    "dirn": arr(v.dirn, 3 * k, np.float64).reshape(k, 3),
                    "area": arr(v.area, k, np.float32), "gravity": arr(v.gravity, k, np.float32),
                    "flux": arr(v.flux, k, np.float32), "ecl_off": arr(v.ecl_off, k + 1, np.int64),
                    "ingress": arr(v.ingress, ne, np.float64) if ne else np.zeros(0),
                    "egress": arr(v.egress, ne, np.float64) if ne else np.zeros(0)}
            pc = FluxParamsC()
            check(L.lch_setup_params(h, ctypes.byref(pc)))
            self.params = _params_from_c(pc)
            self.phase, self.expose, self.slfac = np.zeros(n), np.zeros(n), np.zeros(n)
            self.ndiv = np.zeros(n, dtype=np.int32)
            third = ctypes.c_double()
            check(L.lch_setup_times(h, n, self.phase.ctypes.data_as(_native.D), self.expose.ctypes.data_as(_native.D),
                                    self.ndiv.ctypes.data_as(_native.I), self.slfac.ctypes.data_as(_native.D),
                                    ctypes.byref(third)))
            self.third = third.value
        finally:
            L.lch_setup_free(h)


def prepare(model, times, expose=0.001, ndiv=1):
    """The engine's CPU setup of `model` at `times` (see Setup)."""
    return Setup(model, times, expose, ndiv)


class GridModel:
    """Grids uploaded to the current device (lch_model_upload). `grids` maps
    names of GRID_NAMES to grid dicts (absent = empty); `params` is a dict of
    lch_flux_params fields."""

    def __init__(self, grids, params):
        from . import _ctx
        self._ctx = _ctx()
        keep = []
        arr = (GridSoaC * LCH_NGRID)()
        for name in grids:
            if name not in GRID_NAMES:
                raise ValueError(f"unknown grid '{name}'")
        for gi, name in enumerate(GRID_NAMES):
            g = grids.get(name)
            if g is None:
                continue
            posn = np.ascontiguousarray(g["posn"], dtype=np.float64).reshape(-1)
            dirn = np.ascontiguousarray(g["dirn"], dtype=np.float64).reshape(-1)
            flux = np.ascontiguousarray(g["flux"], dtype=np.float32)
            n = len(flux)
            if len(posn) != 3 * n or len(dirn) != 3 * n:
                raise ValueError(f"grid '{name}': posn and dirn must be [n, 3]")
            c = arr[gi]
            c.n = n
            c.posn = posn.ctypes.data_as(_native.D)
            c.dirn = dirn.ctypes.data_as(_native.D)
            c.flux = flux.ctypes.data_as(_native.F)
            keep += [posn, dirn, flux]
            for k in ("area", "gravity"):
                if g.get(k) is not None:
                    a = np.ascontiguousarray(g[k], dtype=np.float32)
                    setattr(c, k, a.ctypes.data_as(_native.F))
                    keep.append(a)
            if g.get("ecl_off") is not None:
                off = np.ascontiguousarray(g["ecl_off"], dtype=np.int64)
                ing = np.ascontiguousarray(g["ingress"], dtype=np.float64)
                egr = np.ascontiguousarray(g["egress"], dtype=np.float64)
                if len(off) != n + 1 or len(ing) != off[-1] or len(egr) != off[-1]:
                    raise ValueError(f"grid '{name}': ecl_off must be [n + 1] and end at len(ingress)")
                c.ecl_off = off.ctypes.data_as(_native.LL)
                c.ingress = ing.ctypes.data_as(_native.D)
                c.egress = egr.ctypes.data_as(_native.D)
                keep += [off, ing, egr]
        pc = _params_to_c(params)
        h = ctypes.c_void_p()
        check(_native.lib().lch_model_upload(self._ctx, arr, ctypes.byref(pc), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _native.lib().lch_grids_free(h)
            self._h = None

    def flux(self, phase, expose, ndiv, slfac=None, third=0.0):
        """slfac * comp_light(phase, expose, ndiv) + third per time (lch_flux_eval)."""
        ph = np.ascontiguousarray(phase, dtype=np.float64)
        n = len(ph)
        ex = np.ascontiguousarray(np.broadcast_to(expose, (n,)), dtype=np.float64)
        nd = np.ascontiguousarray(np.broadcast_to(ndiv, (n,)), dtype=np.int32)
        sl = np.ascontiguousarray(np.ones(n) if slfac is None else np.broadcast_to(slfac, (n,)), dtype=np.float64)
        out = np.zeros(n)
        check(_native.lib().lch_flux_eval(self._ctx, self._h, ph.ctypes.data_as(_native.D),
                                          ex.ctypes.data_as(_native.D), nd.ctypes.data_as(_native.I),
                                          sl.ctypes.data_as(_native.D), float(third),
                                          out.ctypes.data_as(_native.D), n))
        return out


def upload_setup(setup):
    """GridModel of a Setup's grids and parameters."""
    return GridModel(setup.grids, setup.params)
