"""Learned surrogate: a multilayer perceptron trained and evaluated on the
GPU by the engine's BF16 MFMA kernels (include/lcurve_hip.h, lch_mlp_*).

It maps the five free parameters (q, iangle, r1, t1, t2) to a light curve
folded into phase bins (log of flux / wdwarf), and serves as the cheap
first stage of delayed-acceptance MCMC (lcurve_rs.mcmc). The exact flux path
stays the judge of every accepted sample.
"""
import ctypes

import numpy as np

from . import _ctx, _native
from ._native import check

_F = ctypes.POINTER(ctypes.c_float)
_D = ctypes.POINTER(ctypes.c_double)


def _fp(a):
    return a.ctypes.data_as(_F)


def _dp(a):
    return a.ctypes.data_as(_D)


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None and a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


class Mlp:
    """dims = (in, h1, ..., out): GELU hidden layers, linear output."""

    def __init__(self, dims, seed=1, max_batch=4096):
        self.dims = tuple(int(d) for d in dims)
        d = (ctypes.c_int * len(self.dims))(*self.dims)
        h = ctypes.c_void_p()
        check(_native.lib().lch_mlp_create(_ctx(), d, len(self.dims), int(seed), int(max_batch), ctypes.byref(h)))
        self._h = h
        self.max_batch = int(max_batch)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _native.lib().lch_mlp_free(h)
            self._h = None

    @property
    def n_params(self):
        return int(_native.lib().lch_mlp_param_count(self._h))

    def get_params(self):
        out = np.zeros(self.n_params, dtype=np.float32)
        check(_native.lib().lch_mlp_get_params(self._h, _fp(out)))
        return out

    def set_params(self, p):
        p = _f32(p, (self.n_params,))
        check(_native.lib().lch_mlp_set_params(self._h, _fp(p)))

    def set_data(self, x, y):
        x = _f32(x)
        y = _f32(y)
        if x.ndim != 2 or x.shape[1] != self.dims[0] or y.shape != (x.shape[0], self.dims[-1]):
            raise ValueError("x must be [n, in] and y [n, out]")
        self._data = (x, y)  # keep the host copies alive for the call
        check(_native.lib().lch_mlp_set_data(self._h, _fp(x), _fp(y), x.shape[0]))
        self.n_data = x.shape[0]

    def train_epoch(self, perm, batch, lr):
        perm = np.ascontiguousarray(perm, dtype=np.int32)
        loss = ctypes.c_double()
        check(_native.lib().lch_mlp_train_epoch(self._h, perm.ctypes.data_as(_native.I), len(perm), int(batch),
                                                ctypes.c_float(lr), ctypes.byref(loss)))
        return loss.value

    def loss_grad(self, x, y):
        x = _f32(x)
        y = _f32(y)
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = ctypes.c_double()
        check(_native.lib().lch_mlp_loss_grad(self._h, _fp(x), _fp(y), x.shape[0], ctypes.byref(loss), _fp(g)))
        return loss.value, g

    def adam_step(self, lr):
        check(_native.lib().lch_mlp_adam_step(self._h, ctypes.c_float(lr)))

    def predict(self, x):
        x = _f32(x)
        y = np.zeros((x.shape[0], self.dims[-1]), dtype=np.float32)
        if x.shape[0]:
            check(_native.lib().lch_mlp_predict(self._h, _fp(x), x.shape[0], _fp(y)))
        return y

    def time_step(self, batch, reps=20):
        a, b, f = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check(_native.lib().lch_mlp_time_step(self._h, int(batch), int(reps), ctypes.byref(a), ctypes.byref(b),
                                              ctypes.byref(f)))
        return {"ms_step": a.value, "ms_gemm": b.value, "gemm_flop": f.value}


def n_poly_features(ndim, degree):
    """Columns of the polynomial baseline: 1 and every monomial up to
    ``degree`` (the engine's count)."""
    n = _native.lib().lch_surr_lik_n_features(int(ndim), int(degree))
    if n < 0:
        raise ValueError("bad ndim or degree")
    return n


class SurrogateLikelihood:
    """The autoscaled chi-squared log-likelihood of a time-output surrogate
    against one observed curve, evaluated on the GPU (lch_surr_lik_*):
    curve = exp(poly(u) . coef + y_std * net(u)), u = (p - center) . M.
    Keeps the network alive; retraining the network changes it."""

    def __init__(self, net, center, M, degree, coef, y_std, y, ferr, w):
        center = np.ascontiguousarray(center, dtype=np.float64)
        M = np.ascontiguousarray(M, dtype=np.float64)
        coef = np.ascontiguousarray(coef, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        ferr = np.ascontiguousarray(ferr, dtype=np.float64)
        w = np.ascontiguousarray(w, dtype=np.float64)
        nd = len(center)
        nb = len(y)
        if M.shape != (nd, nd) or coef.shape != (n_poly_features(nd, degree), nb) or ferr.shape != (nb,) \
                or w.shape != (nb,):
            raise ValueError("inconsistent shapes")
        self.net = net
        self.ndim = nd
        h = ctypes.c_void_p()
        check(_native.lib().lch_surr_lik_create(net._h, nd, _dp(center), _dp(M), int(degree), _dp(coef), nb,
                                                ctypes.c_double(y_std), _dp(y), _dp(ferr), _dp(w), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _native.lib().lch_surr_lik_free(h)
            self._h = None

    def __call__(self, params):
        p = np.ascontiguousarray(np.atleast_2d(params), dtype=np.float64)
        if p.shape[1] != self.ndim:
            raise ValueError("parameter rows have the wrong width")
        out = np.zeros(len(p))
        if len(p):
            check(_native.lib().lch_surr_lik_eval(self._h, _dp(p), len(p), _dp(out)))
        return out
