"""Benchmark and test workloads of the flux path.

The three exact-flux workloads, all derived from the reference's test model
(test_data/test_model.dat: white dwarf + Roche-lobe-filling donor):

  wd_donor_45k  — q=0.5, i=82 deg, r1=0.015, Roche-filling star 2, t1=15000 K,
                  t2=3500 K, linear limb darkening 0.4 / 0.5, gravity
                  darkening 0.25 / 0.08, nlat 63 / 177 (5,067 + 39,846
                  elements); 500 times in [-0.2, 1.2], expose 0.001, ndiv 1.
  disc_spot_300k — the same binary plus an opaque flared disc (rdisc 0.1-0.3,
                  T ~ r^-0.75, nrad 155: 100,713 elements), its rim, a bright
                  spot, mutual irradiation and Claret 4-coefficient limb
                  darkening; star 2 at nlat 397 (200,517 elements); 5000
                  times in [-0.5, 1.5], expose 0.005, ndiv 5.
  chisq_4096    — 4096 seeded parameter sets (q~U[0.1,1], i~U[70,90] deg,
                  r1~U[0.005,0.03], t1~U[8000,30000] K, t2~U[2500,5000] K,
                  star-2 nlat ~U{89..344}: 10k-150k elements), each scored
                  against its own curve of 2000 times with 1% Gaussian noise,
                  autoscaled.

``model_text`` edits a model file's values in place, so every workload stays
a plain parameter file the reference's parser accepts.
"""
import math
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

# The reference's test model (tests/golden/test_model.dat holds the same text).
_BASE_MODEL = """q                    = 0.5 0.1 0.01 0
iangle               = 82.0 2.0 0.1 0
r1                   = 0.015 0.005 0.001 0
r2                   = -1 0.01 0.001 0
cphi3                = 0.015 0.001 0.001 0
cphi4                = 0.017 0.001 0.001 0
t1                   = 15000 500 100 0
t2                   = 3500 200 50 0
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
period               = 1.0 1e-06 1e-06 0
gravity_dark1        = 0.25 0.0001 0.0001 0
gravity_dark2        = 0.08 0.0001 0.0001 0
absorb               = 0.5 0.001 0.001 0
slope                = 0 0.01 1e-05 0
quad                 = 0 0.01 1e-05 0
cube                 = 0 0.01 1e-05 0
third                = 0 0.01 1e-05 0
rdisc1               = 0.05 0.01 0.0001 0
rdisc2               = 0.35 0.01 0.02 0
height_disc          = 0.1 0.01 1e-05 0
beta_disc            = 1.5 0.01 1e-05 0
temp_disc            = 3000 50 40 0
texp_disc            = -1.8 0.2 0.001 0
lin_limb_disc        = 0.3 0.02 0.0001 0
quad_limb_disc       = 0 0.02 0.0001 0
radius_spot          = 0.35 0.005 0.01 0
length_spot          = 0.02 0.002 0.005 0
height_spot          = 0.05 0.005 1e-05 0
expon_spot           = 0 0.2 0.1 0
epow_spot            = 1 0.1 0.1 0
angle_spot           = 140 2 2 0
yaw_spot             = 0 2 2 0
temp_spot            = 15000 400 200 0
tilt_spot            = 90 5 2 0
cfrac_spot           = 0.2 0.02 0.008 0
delta_phase          = 1e-07
nlat1f               = 20
nlat2f               = 40
nlat1c               = 10
nlat2c               = 20
npole                = 0
nlatfill             = 0
nlngfill             = 0
lfudge               = 0.05
llo                  = 0
lhi                  = -50
phase1               = 0.05
phase2               = 0.45
wavelength           = 550
roche1               = 0
roche2               = 1
eclipse1             = 1
eclipse2             = 1
glens1               = 0
tperiod              = 0.15
gdark_bolom1         = 1
gdark_bolom2         = 1
mucrit1              = 0
mucrit2              = 0
limb1                = Poly
limb2                = Poly
use_radii            = 1
mirror               = 0
add_disc             = 0
nrad                 = 40
opaque               = 0
add_spot             = 0
nspot                = 100
iscale               = 0
"""


def base_model_text():
    return _BASE_MODEL


def model_text(overrides=None, base=None):
    """Model-file text with values replaced.

    For a physical parameter only the value field changes (range, dstep and
    vary are kept); a parameter absent from the text is appended as
    ``name = value 0 0 0``. Computational parameters take the new value."""
    text = base if base is not None else _BASE_MODEL
    lines = text.splitlines()
    for name, value in (overrides or {}).items():
        pat = re.compile(r"^(\s*" + re.escape(name) + r"\s*=\s*)(\S+)(.*)$")
        for i, ln in enumerate(lines):
            m = pat.match(ln)
            if m:
                lines[i] = f"{m.group(1)}{value}{m.group(3)}"
                break
        else:
            lines.append(f"{name:20s} = {value} 0 0 0" if _is_physical(name) else f"{name:20s} = {value}")
    return "\n".join(lines) + "\n"


_PHYSICAL = set("""q iangle r1 r2 cphi3 cphi4 spin1 spin2 t1 t2 ldc1_1 ldc1_2 ldc1_3 ldc1_4 ldc2_1 ldc2_2
ldc2_3 ldc2_4 velocity_scale beam_factor1 beam_factor2 t0 period pdot deltat gravity_dark1 gravity_dark2
absorb slope quad cube third rdisc1 rdisc2 height_disc beta_disc temp_disc texp_disc lin_limb_disc
quad_limb_disc temp_edge absorb_edge radius_spot length_spot height_spot expon_spot epow_spot angle_spot
yaw_spot temp_spot tilt_spot cfrac_spot""".split())


def _is_physical(name):
    return name in _PHYSICAL or name.startswith("stsp") or name.startswith("uesp")


def linspace_times(t1, t2, n):
    """Times as the reference's lroche / light_curve build them."""
    if n == 1:
        return np.array([(t1 + t2) / 2.0])
    return t1 + (t2 - t1) * np.arange(n, dtype=np.float64) / (n - 1)


# ---------------------------------------------------------------- workloads
WD_DONOR_OVERRIDES = {"ldc2_1": 0.5, "nlat1f": 63, "nlat1c": 63, "nlat2f": 177, "nlat2c": 177}

DISC_SPOT_OVERRIDES = dict(WD_DONOR_OVERRIDES)
DISC_SPOT_OVERRIDES.update({
    "nlat2f": 397, "nlat2c": 397,
    "add_disc": 1, "opaque": 1, "rdisc1": 0.1, "rdisc2": 0.3, "texp_disc": -0.75, "temp_disc": 6000,
    "nrad": 155, "temp_edge": 5000, "absorb_edge": 0.5,
    "add_spot": 1, "nspot": 100, "radius_spot": 0.3, "expon_spot": 1.0, "epow_spot": 1.0,
    "limb1": "Claret", "limb2": "Claret",
    "ldc1_1": 0.55, "ldc1_2": -0.35, "ldc1_3": 0.65, "ldc1_4": -0.25,
    "ldc2_1": 0.70, "ldc2_2": -0.45, "ldc2_3": 0.85, "ldc2_4": -0.35,
})


def wd_donor(scale_nlat=1.0):
    """(model_text, times, expose, ndiv) of wd_donor_45k; ``scale_nlat`` < 1
    shrinks the grids for quick tests."""
    ov = dict(WD_DONOR_OVERRIDES)
    if scale_nlat != 1.0:
        for k in ("nlat1f", "nlat1c", "nlat2f", "nlat2c"):
            ov[k] = max(8, int(round(ov[k] * scale_nlat)))
    return model_text(ov), linspace_times(-0.2, 1.2, 500), 0.001, 1


def disc_spot(scale=1.0, ntime=5000):
    """(model_text, times, expose, ndiv) of disc_spot_300k, optionally with
    grids and times scaled down for tests."""
    ov = dict(DISC_SPOT_OVERRIDES)
    if scale != 1.0:
        for k in ("nlat1f", "nlat1c", "nlat2f", "nlat2c", "nrad"):
            ov[k] = max(8, int(round(ov[k] * scale)))
    return model_text(ov), linspace_times(-0.5, 1.5, ntime), 0.005, 5


# Grids of the surrogate's training set over the disc model's parameter box:
# 200,000 CPU setups at these resolutions take ~0.05 core-s each (the disc's
# radial count dominates the cost; at C2's full grids a set takes seconds).
DISC_BOX_GRIDS = {"nlat1f": 8, "nlat1c": 8, "nlat2f": 40, "nlat2c": 40, "nrad": 16}


def disc_box_text(grids=None):
    """The disc_spot_300k model (white dwarf, Roche-lobe-filling donor,
    opaque flared disc with its rim, bright spot, irradiation, Claret limb
    darkening on both stars) at reduced grids, with ephemeris t0 = 0 and
    period = 1 so that time is phase: the model whose parameter box (q,
    iangle, r1, t1, t2) the surrogate is trained over."""
    ov = dict(DISC_SPOT_OVERRIDES)
    ov.update(DISC_BOX_GRIDS if grids is None else grids)
    ov.update({"t0": 0.0, "period": 1.0, "pdot": 0.0, "deltat": 0.0})
    return model_text(ov)


def _splitmix64(state):
    """splitmix64 stream (Steele, Lea, Flood 2014): portable seeded uniforms."""
    mask = (1 << 64) - 1
    while True:
        state = (state + 0x9E3779B97F4A7C15) & mask
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        z ^= z >> 31
        yield (z >> 11) * (1.0 / (1 << 53))


CHISQ_NAMES = ["q", "iangle", "r1", "t1", "t2", "nlat2"]


def chisq_sets(nsets=4096, seed=42, nlat_lo=89, nlat_hi=344):
    """Seeded parameter sets of chisq_4096: array [nsets, 6] of
    (q, iangle, r1, t1, t2, nlat2)."""
    g = _splitmix64(seed)
    out = np.zeros((nsets, 6))
    for i in range(nsets):
        q = 0.1 + 0.9 * next(g)
        ia = 70.0 + 20.0 * next(g)
        r1 = 0.005 + 0.025 * next(g)
        t1 = 8000.0 + 22000.0 * next(g)
        t2 = 2500.0 + 2500.0 * next(g)
        nl = nlat_lo + min(int(next(g) * (nlat_hi - nlat_lo + 1)), nlat_hi - nlat_lo)
        out[i] = (q, ia, r1, t1, t2, nl)
    return out


def chisq_base_text():
    return model_text(WD_DONOR_OVERRIDES)


def chisq_noise(nsets, ntime, seed=4242):
    """Standard normal deviates [nsets, ntime] for the synthetic observations
    (Box-Muller on the splitmix64 stream)."""
    g = _splitmix64(seed)
    n = nsets * ntime
    u = np.fromiter((next(g) for _ in range(n + (n & 1))), dtype=np.float64, count=n + (n & 1))
    u1 = np.maximum(u[0::2], 1e-300)
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * math.pi * u2), r * np.sin(2 * math.pi * u2)])
    return z[:n].reshape(nsets, ntime)
