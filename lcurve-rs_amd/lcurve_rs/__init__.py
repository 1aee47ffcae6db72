"""lcurve_rs — Python front-end of the MI355X light-curve flux engine.

Mirrors the reference's Python surface (python/src/lib.rs:57-846 and
python/lcurve_rs/__init__.py:29-39): ``Model`` (load / get / set
parameters, ``light_curve``, ``chisq_batch``), ``LcResult`` and the global
device choice. The flux sums run on the GPU through liblcurve_hip.so;
grid and eclipse setup runs on the host CPU with the reference's method.

Differences from the reference, by design:
  * devices are ``"hip"``, ``"hip:N"`` (``"gpu"``, ``"cuda"``, ``"cuda:N"``
    are accepted as aliases so existing scripts keep working);
  * ``"cpu"`` is rejected: this engine has no CPU flux path, and a failing
    GPU call raises instead of silently falling back
    (python/src/lib.rs:784-789 falls back to the CPU);
  * ``Model.light_curve`` also runs on the GPU (always CPU in the reference).
"""
import ctypes
import threading

import numpy as np

from . import _native
from ._native import LcurveError, check

__all__ = ["Model", "LcResult", "Batch", "DevicePackBatch", "set_device", "get_device", "has_hip", "device_count",
           "LcurveError", "read_data"]

_state = {"device": 0}
_ctx_lock = threading.Lock()
_contexts = {}


def device_count():
    """Number of visible HIP devices."""
    return int(_native.lib().lch_device_count())


def has_hip():
    """True when the engine library loads and a HIP device is visible."""
    try:
        return device_count() > 0
    except ImportError:
        return False


def _parse_device(s):
    # python/src/lib.rs:30-50 with HIP device names.
    s = s.strip().lower()
    if s == "cpu":
        raise ValueError("this engine runs the flux on the GPU only; device 'cpu' is not available")
    if s in ("hip", "gpu", "cuda"):
        return 0
    for pre in ("hip:", "cuda:"):
        if s.startswith(pre):
            rest = s[len(pre):]
            try:
                dev = int(rest)
            except ValueError:
                raise ValueError(f"invalid device id: '{rest}'") from None
            if dev < 0:
                raise ValueError(f"device id must be >= 0, got {dev}")
            return dev
    raise ValueError(f"unknown device '{s}': expected 'hip', 'hip:N' (or 'cuda', 'cuda:N')")


def set_device(device):
    """Select the compute device: 'hip' or 'hip:N'."""
    _state["device"] = _parse_device(device)


def get_device():
    """Current device string, 'hip:N'."""
    return f"hip:{_state['device']}"


def _ctx():
    dev = _state["device"]
    with _ctx_lock:
        h = _contexts.get(dev)
        if h is None:
            out = ctypes.c_void_p()
            check(_native.lib().lch_ctx_create(dev, ctypes.byref(out)))
            h = out
            _contexts[dev] = h
        return h


def set_tuning(chunk, persistent=0):
    """Fix the flux-kernel launch form of batches built from now on on the
    current device: `chunk` elements per work item (rounded up to whole
    64-element patches, at most 4096); `persistent` must be 0 (one
    workgroup per item, the only launch form). Batch.autotune() overrides
    it."""
    check(_native.lib().lch_ctx_set_tuning(_ctx(), int(chunk), int(persistent)))


def set_limb(on, max_part=48, check_elements=True):
    """Limb lanes for batches built from now on (default on): no-eclipse
    patches are split per lane into closed-form moments, limb lanes evaluated
    element by element (regrouped across tiles) and lanes that see none of
    the patch, where at most `max_part` of a tile's lanes see it in part.
    Off: every lane of the tile runs the element loop. check_elements: the
    host packer decides the lanes the split's cone test leaves undecided from
    the patch's elements (default); False keeps them limb lanes, the device
    pack's rule (``DevicePackBatch``)."""
    check(_native.lib().lch_ctx_set_limb(_ctx(), int(max_part) if on else 0))
    check(_native.lib().lch_ctx_set_limb_check(_ctx(), int(bool(check_elements))))


def sync():
    """Wait for the work enqueued on the current device's context stream."""
    check(_native.lib().lch_ctx_sync(_ctx()))


def _dptr(a):
    return a.ctypes.data_as(_native.D)


def _iptr(a):
    return a.ctypes.data_as(_native.I)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def read_data(path):
    """Read a data file (time expose flux ferr weight ndiv), types.rs:229-254.

    Returns a dict of numpy arrays; a negative ferr gives weight 0."""
    L = _native.lib()
    n = ctypes.c_int()
    check(L.lch_data_count(path.encode(), ctypes.byref(n)))
    cols = {k: np.zeros(n.value) for k in ("time", "expose", "flux", "ferr", "weight")}
    ndiv = np.zeros(n.value, dtype=np.int32)
    check(L.lch_data_read(path.encode(), n.value, _dptr(cols["time"]), _dptr(cols["expose"]),
                          _dptr(cols["flux"]), _dptr(cols["ferr"]), _dptr(cols["weight"]),
                          _iptr(ndiv)))
    cols["ndiv"] = ndiv
    return cols


class LcResult:
    """Result of ``Model.light_curve`` (python/src/lib.rs:64-116)."""

    def __init__(self, times, calc, res):
        self._times = times
        self._calc = calc
        self.wdwarf = res.wdwarf
        self.chisq = res.chisq
        self.wnok = res.wnok
        self.logg1 = res.logg1
        self.logg2 = res.logg2
        self.rv1 = res.rv1
        self.rv2 = res.rv2
        self._sfac = np.array(list(res.sfac))

    @property
    def calc(self):
        return self._calc.copy()

    @property
    def flux(self):
        return self._calc.copy()

    @property
    def times(self):
        return self._times.copy()

    @property
    def sfac(self):
        return self._sfac.copy()

    def __repr__(self):
        return (f"LcResult(npoints={len(self._calc)}, wdwarf={self.wdwarf:.6e}, "
                f"chisq={self.chisq:.4f}, logg1={self.logg1:.4f}, logg2={self.logg2:.4f})")


_WHITEDWARF_TEMPLATE = """q                    = {q} 0.1 0.01 0
iangle               = {iangle} 2.0 0.1 0
r1                   = {r1} 0.005 0.001 0
r2                   = -1 0.01 0.001 0
cphi3                = 0.015 0.001 0.001 0
cphi4                = 0.017 0.001 0.001 0
t1                   = {t1} 500 100 0
t2                   = {t2} 200 50 0
spin1                = 1 0.001 0.001 0
spin2                = 1 0.001 0.001 0
ldc1_1               = 0.4 0.01 0.01 0
ldc1_2               = 0.0 0.01 0.01 0
ldc1_3               = 0.0 0.01 0.01 0
ldc1_4               = 0.0 0.01 0.01 0
ldc2_1               = 0.6 0.01 0.01 0
ldc2_2               = 0.0 0.01 0.01 0
ldc2_3               = 0.0 0.01 0.01 0
ldc2_4               = 0.0 0.01 0.01 0
velocity_scale       = 0 1 1 0
beam_factor1         = 0 0.1 0.02 0
beam_factor2         = 0 0.1 0.002 0
deltat               = 0 0.001 0.001 0
t0                   = 0.0 0.0001 1e-05 0
period               = {period} 1e-06 1e-06 0
gravity_dark1        = 0.25 0.0001 0.0001 0
gravity_dark2        = 0.08 0.0001 0.0001 0
absorb               = 1.0 0.001 0.001 0
slope                = 0 0.001 0.0001 0
quad                 = 0 0.001 0.0001 0
cube                 = 0 0.001 0.0001 0
third                = 0 0.001 0.0001 0
rdisc1               = {disc_r1} 0.1 0.01 0
rdisc2               = {disc_r2} 0.1 0.01 0
height_disc          = 0.02 0.01 0.001 0
beta_disc            = 2.0 0.1 0.1 0
temp_disc            = {disc_temp} 500 100 0
texp_disc            = -0.75 0.1 0.01 0
lin_limb_disc        = 0.3 0.1 0.01 0
quad_limb_disc       = 0.0 0.1 0.01 0
temp_edge            = 0.0 500 100 0
absorb_edge          = 0.0 0.001 0.001 0
radius_spot          = 0.0 0.001 0.001 0
length_spot          = 0.0 0.001 0.001 0
height_spot          = 0.0 0.001 0.001 0
expon_spot           = 0.0 0.001 0.001 0
epow_spot            = 1.0 0.001 0.001 0
angle_spot           = 0.0 0.001 0.001 0
yaw_spot             = 0.0 0.001 0.001 0
temp_spot            = 0.0 0.001 0.001 0
tilt_spot            = 90.0 0.001 0.001 0
cfrac_spot           = 0.0 0.001 0.001 0
delta_phase          = -0.001
nlat1f               = 50
nlat2f               = 50
nlat1c               = 50
nlat2c               = 50
npole                = 0
nlatfill             = 4
nlngfill             = 4
lfudge               = 0.05
llo                  = 0.0
lhi                  = 0.0
wavelength           = 4700
roche1               = 1
roche2               = 1
eclipse1             = 1
eclipse2             = 1
glens1               = 0
use_radii            = 0
tperiod              = 0
gdark_bolom1         = 0
gdark_bolom2         = 0
mucrit1              = 0.0
mucrit2              = 0.0
limb1                = Claret
limb2                = Claret
mirror               = 0
add_disc             = {add_disc}
opaque               = 1
add_spot             = 0
nspot                = 0
nrad                 = 50
phase1               = -0.2
phase2               = 1.2
iscale               = 1
"""


class Model:
    """An lcurve model (python/src/lib.rs:255-744)."""

    def __init__(self, path=None, *, _handle=None):
        L = _native.lib()
        if _handle is not None:
            self._h = _handle
        else:
            if path is None:
                raise TypeError("Model(path) needs a parameter file path")
            out = ctypes.c_void_p()
            check(L.lch_model_load_file(str(path).encode(), ctypes.byref(out)))
            self._h = out

    @classmethod
    def from_text(cls, text):
        """Parse model-file text held in memory."""
        out = ctypes.c_void_p()
        check(_native.lib().lch_model_load_text(text.encode(), ctypes.byref(out)))
        return cls(_handle=out)

    @staticmethod
    def whitedwarf(q=0.5, iangle=82.0, t1=15000.0, t2=3500.0, period=0.1, r1=0.015, disc=False):
        """CV model with the reference's defaults (python/src/lib.rs:414-562)."""
        text = _WHITEDWARF_TEMPLATE.format(
            q=q, iangle=iangle, t1=t1, t2=t2, period=period, r1=r1,
            disc_r1=0.2 if disc else 0.0, disc_r2=0.45 if disc else 0.0,
            disc_temp=8000.0 if disc else 0.0, add_disc=1 if disc else 0)
        return Model.from_text(text)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _native._lib is not None:
            _native._lib.lch_model_free(h)
            self._h = None

    def get_param(self, name):
        v = ctypes.c_double()
        check(_native.lib().lch_model_get_param(self._h, name.encode(), ctypes.byref(v)))
        return v.value

    def set_param(self, name, value):
        check(_native.lib().lch_model_set_param(self._h, name.encode(), float(value)))

    @property
    def iscale(self):
        """Whether the curve is fitted per component (model.rs:181)."""
        v = ctypes.c_int()
        check(_native.lib().lch_model_iscale(self._h, ctypes.byref(v)))
        return bool(v.value)

    def get_pparam(self, name):
        v, r, d = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        vary, defined = ctypes.c_int(), ctypes.c_int()
        check(_native.lib().lch_model_get_pparam(self._h, name.encode(), ctypes.byref(v), ctypes.byref(r),
                                                 ctypes.byref(d), ctypes.byref(vary), ctypes.byref(defined)))
        return {"value": v.value, "range": r.value, "dstep": d.value, "vary": bool(vary.value),
                "defined": bool(defined.value)}

    def set_pparam(self, name, *, value=None, range=None, dstep=None, vary=None):
        def opt(x, ct):
            return ctypes.byref(ct(x)) if x is not None else None
        check(_native.lib().lch_model_set_pparam(self._h, name.encode(), opt(value, ctypes.c_double),
                                                 opt(range, ctypes.c_double), opt(dstep, ctypes.c_double),
                                                 opt(None if vary is None else int(bool(vary)), ctypes.c_int)))

    def get_free_params(self):
        L = _native.lib()
        out = []
        for i in range(L.lch_param_count()):
            name = L.lch_param_name(i).decode()
            p = self.get_pparam(name)
            if p["vary"] and p["defined"]:
                out.append({"name": name, "value": p["value"], "range": p["range"], "dstep": p["dstep"]})
        return out

    def set_params(self, names, values):
        names = list(names)
        values = list(values)
        if len(names) != len(values):
            raise ValueError(f"names and values must have same length ({len(names)} vs {len(values)})")
        for n, v in zip(names, values):
            self.set_param(n, v)

    def copy(self):
        out = ctypes.c_void_p()
        check(_native.lib().lch_model_copy(self._h, ctypes.byref(out)))
        return Model(_handle=out)

    def _prop(name):
        return property(lambda self: self.get_param(name), lambda self, v: self.set_param(name, v))

    q = _prop("q")
    iangle = _prop("iangle")
    r1 = _prop("r1")
    r2 = _prop("r2")
    t1 = _prop("t1")
    t2 = _prop("t2")
    period = _prop("period")
    t0 = _prop("t0")
    velocity_scale = _prop("velocity_scale")
    del _prop

    def light_curve(self, *, time1=-0.2, time2=1.2, ntime=500, times=None, expose=0.001, ndiv=1,
                    data=None, scale=False):
        """Light curve on the GPU (python/src/lib.rs:617-676 semantics)."""
        if data is not None:
            cols = read_data(data)
            rdata = True
        else:
            if times is not None:
                t = _f64(times)
            else:
                if ntime == 0:
                    raise ValueError("ntime must be > 0")
                if ntime == 1:
                    t = np.array([(time1 + time2) / 2.0])
                else:
                    t = time1 + (time2 - time1) * np.arange(ntime, dtype=np.float64) / (ntime - 1)
            n = len(t)
            cols = {"time": t, "expose": np.full(n, float(expose)), "flux": np.zeros(n),
                    "ferr": np.ones(n), "weight": np.zeros(n), "ndiv": np.full(n, int(ndiv), dtype=np.int32)}
            rdata = False
        n = len(cols["time"])
        calc = np.zeros(n)
        res = _native.LcResultC()
        check(_native.lib().lch_light_curve(
            _ctx(), self._h, _dptr(cols["time"]), _dptr(cols["expose"]), _dptr(cols["flux"]),
            _dptr(cols["ferr"]), _dptr(cols["weight"]), _iptr(cols["ndiv"]), n, int(bool(scale)),
            int(rdata), _dptr(calc), ctypes.byref(res)))
        return LcResult(cols["time"].copy(), calc, res)

    def selfcheck_patch_codes(self, times, expose=0.001, ndiv=1):
        """CPU check (no device) of the engine's work skipping for this model
        at these times: returns (max relative difference between the coded
        and the brute-force FP64 sums, {code: (patch, tile) pairs, and the
        per-lane split of no-eclipse pairs: lanes seen whole / in part /
        not at all})."""
        t = _f64(times)
        n = len(t)
        e = np.full(n, float(expose))
        nd = np.full(n, int(ndiv), dtype=np.int32)
        r = ctypes.c_double()
        c = (ctypes.c_longlong * 4)()
        sp = (ctypes.c_longlong * 3)()
        check(_native.lib().lch_selfcheck_patch_codes(self._h, _dptr(t), _dptr(e), _iptr(nd), n,
                                                      ctypes.byref(r), c, sp))
        return r.value, {"full": c[0], "no_eclipse": c[1], "skip": c[2], "moments": c[3],
                         "split_whole": sp[0], "split_part": sp[1], "split_none": sp[2]}

    def selfcheck_eclipse_prescreen(self):
        """CPU check (no device): both stars' grids built with and without the
        certified setup short cuts (eclipse phase scan pre-screen, Taylor-
        decided surface chop). Returns a dict with
        elements compared, mismatches (must be 0), scan steps skipped and
        evaluated, and the build times in ms without and with it."""
        out = (ctypes.c_longlong * 4)()
        ms = np.zeros(2)
        check(_native.lib().lch_selfcheck_eclipse_prescreen(self._h, out, _dptr(ms)))
        return {"elements": out[0], "mismatches": out[1], "skipped": out[2], "evaluated": out[3],
                "ms_plain": float(ms[0]), "ms_screened": float(ms[1])}

    def chisq_batch(self, data, names, values, scale=True):
        """Chi-squared of N parameter sets (python/src/lib.rs:696-734).

        ``values`` has shape (N, len(names)); unphysical sets give NaN."""
        cols = read_data(data)
        vals = _f64(values)
        if vals.ndim != 2:
            raise ValueError("values must be a 2-D array")
        nsets, ncols = vals.shape
        if ncols != len(names):
            raise ValueError(f"values has {ncols} columns but {len(names)} names were given")
        if len(names) == 0:
            raise ValueError("names must not be empty")
        out = np.zeros(nsets)
        if nsets == 0:
            return out
        carr = (ctypes.c_char_p * len(names))(*[n.encode() for n in names])
        check(_native.lib().lch_chisq_batch(
            _ctx(), self._h, _dptr(cols["time"]), _dptr(cols["expose"]), _dptr(cols["flux"]),
            _dptr(cols["ferr"]), _dptr(cols["weight"]), _iptr(cols["ndiv"]), len(cols["time"]),
            carr, len(names), _dptr(vals), nsets, int(bool(scale)), _dptr(out)))
        return out

    def __repr__(self):
        return (f"Model(q={self.q}, iangle={self.iangle}, r1={self.r1}, r2={self.r2}, "
                f"t1={self.t1}, t2={self.t2})")


class DevicePackBatch:
    """A batch of parameter sets whose per-set pack runs on the GPU
    (lch_dpack_*): the constructor does the CPU setup and each set's host
    layout and stages the raw element grids; every ``run`` / ``time`` packs
    all sets on the GPU from those grids (element records, patch cones and
    moments, work codes, per-lane splits, work items, limb tables) and runs
    the flux pass and the assembly. Results equal a ``Batch`` of the same
    models built with the same chunk and limb threshold (``limb``: 0 off,
    else ``set_limb(True, limb)``), bit for bit. The stand-in for the
    reference's chisq_batch_gpu (lcurve-cuda/src/batch.rs:73-136) after
    prepare_model."""

    def __init__(self, models, times, expose, ndiv, *, flux=None, ferr=None, weight=None, flux2d=None,
                 ferr2d=None, weight2d=None, chunk=2048, limb=0, slots=1):
        L = _native.lib()
        t = _f64(times)
        n = len(t)
        self.n_times = n
        self.n_sets = len(models)
        ex = _f64(np.broadcast_to(expose, (n,)))
        nd = np.ascontiguousarray(np.broadcast_to(ndiv, (n,)), dtype=np.int32)
        fl = _f64(np.zeros(n) if flux is None else flux)
        fe = _f64(np.ones(n) if ferr is None else ferr)
        wt = _f64(np.ones(n) if weight is None else weight)
        f2, e2, w2 = (None if a is None else _f64(a) for a in (flux2d, ferr2d, weight2d))
        harr = (ctypes.c_void_p * self.n_sets)(*[m._h.value for m in models])
        out = ctypes.c_void_p()
        check(L.lch_dpack_build(_ctx(), harr, self.n_sets, _dptr(t), _dptr(ex), _dptr(fl), _dptr(fe), _dptr(wt),
                                _iptr(nd), n, *[None if a is None else _dptr(a) for a in (f2, e2, w2)],
                                int(chunk), int(limb), ctypes.byref(out)))
        self._h = out
        if slots != 1:
            self.set_slots(slots)

    def set_slots(self, slots):
        """Work-table sets launches alternate between (1 or 2): with 2 the
        pack of one launch runs beside the flux pass of the one before."""
        check(_native.lib().lch_dpack_set_slots(_ctx(), self._h, int(slots)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _native._lib is not None:
            _native._lib.lch_dpack_free(h)
            self._h = None

    def run(self, scale=True, want_flux=True):
        """One pack + pass; dict with chisq, wnok, sfac, wdwarf (per set)
        and flux [n_sets, n_times]."""
        ns, nt = self.n_sets, self.n_times
        chisq, wnok, sfac, wd = (np.zeros(ns) for _ in range(4))
        flux = np.zeros((ns, nt)) if want_flux else None
        check(_native.lib().lch_dpack_run(_ctx(), self._h, int(bool(scale)), _dptr(chisq), _dptr(wnok),
                                          _dptr(sfac), _dptr(wd), None if flux is None else _dptr(flux)))
        return {"chisq": chisq, "wnok": wnok, "sfac": sfac, "wdwarf": wd, "flux": flux}

    def launch(self, scale=True):
        check(_native.lib().lch_dpack_launch(_ctx(), self._h, int(bool(scale))))

    def upload(self):
        """Copy the staged grids and tables to the device again."""
        check(_native.lib().lch_dpack_upload(_ctx(), self._h))

    def sync(self):
        sync()

    def time(self, scale=True, reps=5):
        """Mean ms per launch of the pack alone, the flux pass alone and the
        whole launch."""
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check(_native.lib().lch_dpack_time(_ctx(), self._h, int(bool(scale)), int(reps),
                                           ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def set_obs(self, flux2d=None, ferr2d=None, weight2d=None):
        """Replace the observations of all sets ([n_sets, n_times] arrays)."""
        shp = (self.n_sets, self.n_times)
        arrs = [None if a is None else np.ascontiguousarray(np.broadcast_to(a, shp), dtype=np.float64)
                for a in (flux2d, ferr2d, weight2d)]
        check(_native.lib().lch_dpack_set_obs(_ctx(), self._h, *[None if a is None else _dptr(a) for a in arrs]))

    def stats(self):
        s = _native.DPackStatsC()
        check(_native.lib().lch_dpack_stats_get(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def check(self, set_index):
        """Host pack of one set against a fresh device pack: dict of
        mismatch counts (records bytes, cones, moments, codes, segment
        flags, items, item moments and lane-mask groups, limb tables) and
        the host's item and limb item counts."""
        d = (ctypes.c_longlong * 10)()
        check(_native.lib().lch_dpack_check(_ctx(), self._h, int(set_index), d))
        keys = ("record_bytes", "cones", "moments", "codes", "seg_flags", "items", "item_moments", "host_items",
                "limb", "host_limb_items")
        return dict(zip(keys, (int(v) for v in d)))


class Batch:
    """A resident batch of parameter sets (lch_batch_*): CPU setup and the
    upload happen once in the constructor, then ``run`` / ``time`` launch
    the GPU pass over data already in HBM."""

    def __init__(self, model, times, expose, ndiv, names=(), values=None, *, flux=None, ferr=None,
                 weight=None, flux2d=None, ferr2d=None, weight2d=None, keep_maps=False):
        """``model`` is either one Model whose physical parameters ``names``
        take the rows of ``values`` (one set per row), or a list of Models
        (one set each; ``names`` / ``values`` unused)."""
        L = _native.lib()
        t = _f64(times)
        n = len(t)
        self.n_times = n
        ex = _f64(np.broadcast_to(expose, (n,)))
        nd = np.ascontiguousarray(np.broadcast_to(ndiv, (n,)), dtype=np.int32)
        fl = _f64(np.zeros(n) if flux is None else flux)
        fe = _f64(np.ones(n) if ferr is None else ferr)
        wt = _f64(np.ones(n) if weight is None else weight)
        f2 = None if flux2d is None else _f64(flux2d)
        e2 = None if ferr2d is None else _f64(ferr2d)
        w2 = None if weight2d is None else _f64(weight2d)
        flags = _native.LCH_BATCH_KEEP_MAPS if keep_maps else 0
        out = ctypes.c_void_p()
        obs2 = (None if f2 is None else _dptr(f2), None if e2 is None else _dptr(e2),
                None if w2 is None else _dptr(w2))
        if isinstance(model, (list, tuple)):
            self.n_sets = len(model)
            harr = (ctypes.c_void_p * self.n_sets)(*[m._h.value for m in model])
            self._keep = (t, ex, nd, fl, fe, wt, f2, e2, w2, list(model), harr)
            check(L.lch_batch_build_models(
                _ctx(), harr, self.n_sets, _dptr(t), _dptr(ex), _dptr(fl), _dptr(fe), _dptr(wt),
                _iptr(nd), n, *obs2, flags, ctypes.byref(out)))
        else:
            vals = _f64(values if values is not None else np.zeros((1, 0)))
            if vals.ndim == 1:
                vals = vals.reshape(-1, len(names)) if len(names) else vals.reshape(-1, 0)
            self.n_sets = vals.shape[0]
            carr = (ctypes.c_char_p * max(len(names), 1))(*[s.encode() for s in names])
            self._keep = (t, ex, nd, fl, fe, wt, vals, f2, e2, w2, carr)
            check(L.lch_batch_build(
                _ctx(), model._h, _dptr(t), _dptr(ex), _dptr(fl), _dptr(fe), _dptr(wt), _iptr(nd), n,
                carr, len(names), _dptr(vals), self.n_sets, *obs2, flags, ctypes.byref(out)))
        self._h = out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _native._lib is not None:
            _native._lib.lch_batch_free(h)
            self._h = None

    def run(self, scale=True, want_flux=True):
        """One pass; returns dict with chisq, wnok, sfac, wdwarf (per set),
        sfac5 [n_sets, 5] (per-component factors for iscale sets) and
        flux [n_sets, n_times]."""
        ns, nt = self.n_sets, self.n_times
        chisq, wnok, sfac, wd = (np.zeros(ns) for _ in range(4))
        flux = np.zeros((ns, nt)) if want_flux else None
        check(_native.lib().lch_batch_run(_ctx(), self._h, int(bool(scale)), _dptr(chisq), _dptr(wnok),
                                          _dptr(sfac), _dptr(wd), None if flux is None else _dptr(flux)))
        sfac5 = np.zeros((ns, 5))
        if ns:
            check(_native.lib().lch_batch_sfac(self._h, _dptr(sfac5)))
        return {"chisq": chisq, "wnok": wnok, "sfac": sfac, "wdwarf": wd, "sfac5": sfac5, "flux": flux}

    def set_obs(self, flux2d=None, ferr2d=None, weight2d=None):
        """Replace the observations of all sets ([n_sets, n_times] arrays)."""
        shp = (self.n_sets, self.n_times)
        arrs = [None if a is None else np.ascontiguousarray(np.broadcast_to(a, shp), dtype=np.float64)
                for a in (flux2d, ferr2d, weight2d)]
        check(_native.lib().lch_batch_set_obs(_ctx(), self._h, *[None if a is None else _dptr(a) for a in arrs]))

    def launch(self, scale=True):
        check(_native.lib().lch_batch_launch(_ctx(), self._h, int(bool(scale))))

    def sync(self):
        check(_native.lib().lch_ctx_sync(_ctx()))

    def time(self, scale=True, reps=5):
        """Mean ms per pass of the flux kernel and of the whole pass."""
        a, b = ctypes.c_double(), ctypes.c_double()
        check(_native.lib().lch_batch_time(_ctx(), self._h, int(bool(scale)), int(reps),
                                           ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def autotune(self, scale=True):
        check(_native.lib().lch_batch_autotune(_ctx(), self._h, int(bool(scale))))

    def stats(self):
        s = _native.BatchStatsC()
        check(_native.lib().lch_batch_stats_get(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def set_info(self, i):
        valid = ctypes.c_int()
        vals = [ctypes.c_double() for _ in range(4)]
        err = ctypes.create_string_buffer(512)
        check(_native.lib().lch_batch_set_info(self._h, i, ctypes.byref(valid), *[ctypes.byref(v) for v in vals],
                                               err, 512))
        return {"valid": bool(valid.value), "logg1": vals[0].value, "logg2": vals[1].value,
                "rv1": vals[2].value, "rv2": vals[3].value, "error": err.value.decode()}

    def visibility(self, set_index, grid, phases):
        """GPU eclipse visibility mask [nphase, nelem] (needs keep_maps)."""
        ph = _f64(phases)
        n = ctypes.c_int()
        L = _native.lib()
        check(L.lch_batch_visibility(_ctx(), self._h, set_index, grid, _dptr(ph), len(ph), None, ctypes.byref(n)))
        mask = np.zeros((len(ph), n.value), dtype=np.uint8)
        if n.value and len(ph):
            check(L.lch_batch_visibility(_ctx(), self._h, set_index, grid, _dptr(ph), len(ph),
                                         mask.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)), ctypes.byref(n)))
        return mask
