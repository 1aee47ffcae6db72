"""ORACLE — test infrastructure only.

ctypes wrapper of oracle/_build/liblcurve_oracle.so, the FP64 CPU
restatement of the lcurve-rs flux path (see lcurve_oracle.cpp for the
reference file:line map and what pins each part). Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg use it, as the
checker; the product path never does.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# LCURVE_ORACLE_LIB: another build of the same sources (the sanitizer build).
LIB_PATH = os.environ.get("LCURVE_ORACLE_LIB", os.path.join(HERE, "_build", "liblcurve_oracle.so"))

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"oracle not built: {LIB_PATH} (run `make -C oracle`)")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_light_curve.restype = ctypes.c_int
        L.orc_light_curve.argtypes = [ctypes.c_char_p, _D, _D, _D, _D, _D, _I, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int, _D, _D, ctypes.c_char_p, ctypes.c_int]
        L.orc_components.restype = ctypes.c_int
        L.orc_components.argtypes = [ctypes.c_char_p, _D, _D, _D, _D, _D, _I, ctypes.c_int, ctypes.c_int, _D,
                                     ctypes.c_char_p, ctypes.c_int]
        L.orc_batch_create.restype = ctypes.c_void_p
        L.orc_batch_create.argtypes = [ctypes.c_char_p, _D, _D, _D, _D, _D, _I, ctypes.c_int, ctypes.c_char_p,
                                       _D, ctypes.c_int, _D, _D, _D, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
        L.orc_batch_setup_seconds.restype = ctypes.c_double
        L.orc_batch_setup_seconds.argtypes = [ctypes.c_void_p]
        L.orc_batch_flux.restype = ctypes.c_int
        L.orc_batch_flux.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _D, _D]
        L.orc_batch_free.restype = None
        L.orc_batch_free.argtypes = [ctypes.c_void_p]
        L.orc_batch_visibility.restype = ctypes.c_int
        L.orc_batch_visibility.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _D, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_ubyte), ctypes.POINTER(ctypes.c_longlong)]
        L.orc_batch_grid.restype = ctypes.c_longlong
        L.orc_batch_grid.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _D, _D, _D, _D, _I]
        L.orc_phases.restype = ctypes.c_int
        L.orc_phases.argtypes = [ctypes.c_char_p, _D, ctypes.c_int, _D]
        L.orc_lensing_mag.restype = ctypes.c_double
        L.orc_lensing_mag.argtypes = [_D, _D, ctypes.c_double]
        L.orc_batch_gravity.restype = ctypes.c_longlong
        L.orc_batch_gravity.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _D]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_D)


def _f(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None:
        a = np.ascontiguousarray(np.broadcast_to(a, (n,)), dtype=np.float64)
    return a


SCALAR_KEYS = ("wdwarf", "chisq", "wnok", "logg1", "logg2", "rv1", "rv2")


def light_curve(model_text, times, expose=0.001, ndiv=1, flux=None, ferr=None, weight=None,
                scale=False, rdata=False, nthreads=1):
    """light_curve_comp (orchestration.rs:36-318) in FP64 on the CPU.

    Returns (calc, scalars) with scalars holding wdwarf, chisq, wnok, logg1,
    logg2, rv1, rv2, sfac (5) and the setup / flux wall times."""
    t = _f(times)
    n = len(t)
    ex = _f(expose, n)
    nd = np.ascontiguousarray(np.broadcast_to(ndiv, (n,)), dtype=np.int32)
    fl = _f(np.zeros(n) if flux is None else flux, n)
    fe = _f(np.zeros(n) if ferr is None else ferr, n)
    wt = _f(np.ones(n) if weight is None else weight, n)
    calc = np.zeros(n)
    sc = np.zeros(14)
    err = ctypes.create_string_buffer(1024)
    rc = lib().orc_light_curve(model_text.encode(), _d(t), _d(ex), _d(fl), _d(fe), _d(wt),
                               nd.ctypes.data_as(_I), n, int(scale), int(rdata), int(nthreads),
                               _d(calc), _d(sc), err, 1024)
    if rc != 0:
        raise ValueError(err.value.decode())
    out = dict(zip(SCALAR_KEYS, sc[:7]))
    out["sfac"] = sc[7:12].copy()
    out["t_setup"] = sc[12]
    out["t_flux"] = sc[13]
    return calc, out


def components(model_text, times, expose=0.001, ndiv=1, flux=None, ferr=None, weight=None,
               nthreads=1):
    """The iscale columns (orchestration.rs:181-227) of a model, unscaled:
    returns (fcomp [n, 5], ncol); column 4 is zero when t2 <= 0."""
    t = _f(times)
    n = len(t)
    ex = _f(expose, n)
    nd = np.ascontiguousarray(np.broadcast_to(ndiv, (n,)), dtype=np.int32)
    fl = _f(np.zeros(n) if flux is None else flux, n)
    fe = _f(np.ones(n) if ferr is None else ferr, n)
    wt = _f(np.ones(n) if weight is None else weight, n)
    out = np.zeros((n, 5))
    err = ctypes.create_string_buffer(1024)
    ncol = lib().orc_components(model_text.encode(), _d(t), _d(ex), _d(fl), _d(fe), _d(wt),
                                nd.ctypes.data_as(_I), n, int(nthreads), _d(out), err, 1024)
    if ncol < 0:
        raise ValueError(err.value.decode())
    return out, ncol


class OracleBatch:
    """chisq_batch (orchestration.rs:329-357) split into setup and flux.

    ``model_text`` is one model text (parameter sets from ``names`` /
    ``values``) or a list of texts, one per set."""

    def __init__(self, model_text, times, expose, ndiv, names=(), values=None, flux=None, ferr=None,
                 weight=None, flux2d=None, ferr2d=None, weight2d=None, nthreads=1):
        t = _f(times)
        n = len(t)
        self.n = n
        ex = _f(expose, n)
        nd = np.ascontiguousarray(np.broadcast_to(ndiv, (n,)), dtype=np.int32)
        fl = _f(np.zeros(n) if flux is None else flux, n)
        fe = _f(np.ones(n) if ferr is None else ferr, n)
        wt = _f(np.ones(n) if weight is None else weight, n)
        f2 = None if flux2d is None else _f(flux2d)
        e2 = None if ferr2d is None else _f(ferr2d)
        w2 = None if weight2d is None else _f(weight2d)
        err = ctypes.create_string_buffer(1024)
        obs2 = (None if f2 is None else _d(f2), None if e2 is None else _d(e2),
                None if w2 is None else _d(w2))
        if isinstance(model_text, (list, tuple)):
            self.nsets = len(model_text)
            arr = (ctypes.c_char_p * self.nsets)(*[s.encode() for s in model_text])
            self._keep = (t, ex, nd, fl, fe, wt, f2, e2, w2, arr)
            L = lib()
            L.orc_batch_create_texts.restype = ctypes.c_void_p
            L.orc_batch_create_texts.argtypes = [ctypes.POINTER(ctypes.c_char_p), ctypes.c_int, _D, _D, _D, _D,
                                                 _D, _I, ctypes.c_int, _D, _D, _D, ctypes.c_int,
                                                 ctypes.c_char_p, ctypes.c_int]
            h = L.orc_batch_create_texts(arr, self.nsets, _d(t), _d(ex), _d(fl), _d(fe), _d(wt),
                                         nd.ctypes.data_as(_I), n, *obs2, int(nthreads), err, 1024)
        else:
            if len(names):
                vals = _f(values).reshape(-1, len(names))
            else:
                rows = 1 if values is None else max(1, len(values))
                vals = np.zeros((rows, 0))
            self.nsets = vals.shape[0]
            self._keep = (t, ex, nd, fl, fe, wt, vals, f2, e2, w2)
            h = lib().orc_batch_create(model_text.encode(), _d(t), _d(ex), _d(fl), _d(fe), _d(wt),
                                       nd.ctypes.data_as(_I), n, "\n".join(names).encode(), _d(vals),
                                       self.nsets, *obs2, int(nthreads), err, 1024)
        if not h:
            raise ValueError(err.value.decode())
        self._h = h
        self.setup_seconds = lib().orc_batch_setup_seconds(h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_batch_free(h)
            self._h = None

    def flux(self, scale=True, nthreads=1, want_flux=True):
        chisq = np.zeros(self.nsets)
        fl = np.zeros((self.nsets, self.n)) if want_flux else None
        lib().orc_batch_flux(self._h, int(scale), int(nthreads), _d(chisq), None if fl is None else _d(fl))
        return chisq, fl

    def grid(self, set_index, grid):
        """Element geometry of one grid: dict of posn, dirn [n,3], area,
        flux [n] and eclipse-interval counts."""
        n = lib().orc_batch_grid(self._h, set_index, grid, None, None, None, None, None)
        if n < 0:
            raise ValueError("bad set or grid")
        posn, dirn = np.zeros((n, 3)), np.zeros((n, 3))
        area, flux = np.zeros(n), np.zeros(n)
        necl = np.zeros(n, dtype=np.int32)
        lib().orc_batch_grid(self._h, set_index, grid, _d(posn), _d(dirn), _d(area), _d(flux),
                             necl.ctypes.data_as(_I))
        return {"posn": posn, "dirn": dirn, "area": area, "flux": flux, "necl": necl}

    def gravity(self, set_index, grid):
        """Normalised surface gravity (Point::gravity) of grid 0 (star 1,
        fine) or 2 (star 2, fine)."""
        n = lib().orc_batch_gravity(self._h, set_index, grid, None)
        if n < 0:
            raise ValueError("bad set or grid")
        g = np.zeros(n)
        lib().orc_batch_gravity(self._h, set_index, grid, _d(g))
        return g

    def visibility(self, set_index, grid, phases):
        ph = _f(phases)
        ne = ctypes.c_longlong()
        lib().orc_batch_visibility(self._h, set_index, grid, _d(ph), len(ph), None, ctypes.byref(ne))
        mask = np.zeros((len(ph), ne.value), dtype=np.uint8)
        if ne.value and len(ph):
            rc = lib().orc_batch_visibility(self._h, set_index, grid, _d(ph), len(ph),
                                            mask.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)),
                                            ctypes.byref(ne))
            if rc != 0:
                raise ValueError("oracle visibility failed")
        return mask


def phases(model_text, times):
    t = _f(times)
    out = np.zeros(len(t))
    if lib().orc_phases(model_text.encode(), _d(t), len(t), _d(out)) != 0:
        raise ValueError("bad model")
    return out


def lensing_mag(posn, earth, rlens1):
    """lensing_mag (flux.rs:390-406) of one element position and earth vector."""
    p = _f(posn)
    e = _f(earth)
    return lib().orc_lensing_mag(_d(p), _d(e), float(rlens1))
