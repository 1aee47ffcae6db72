"""ORACLE — test infrastructure only.

numpy restatement of the surrogate MLP of lcurve-rs_amd/csrc/surrogate
(no reference counterpart: the surrogate is this build's own addition).
Layers: GELU (erf form) after every hidden layer, linear output, MSE loss
mean((y - t)^2), Adam (Kingma & Ba 2015).

Two modes:
  * bf16=False: exact FP64 forward / backward. Pinned by central finite
    differences (tests/test_mlp_oracle.py).
  * bf16=True: the rounding points of the GPU path. Inputs, hidden
    activations, weights for the GEMMs and the backward GEMM inputs are
    rounded to BF16 (round to nearest even); sums are FP64 here (FP32 on
    the GPU). The GPU is compared with this mode.

Parameters are a flat FP32 vector ordered W0 [out x in], b0, W1, b1, ...
"""
import math

import numpy as np


def bf16(x):
    """Round to BF16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, u | 0x400000, r).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


_erf = np.vectorize(math.erf, otypes=[np.float64])


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def unpack(params, dims):
    out, o = [], 0
    for i in range(len(dims) - 1):
        nw = dims[i + 1] * dims[i]
        W = np.asarray(params[o:o + nw], dtype=np.float64).reshape(dims[i + 1], dims[i])
        o += nw
        b = np.asarray(params[o:o + dims[i + 1]], dtype=np.float64)
        o += dims[i + 1]
        out.append((W, b))
    return out


def forward(params, dims, x, bf16_mode=True):
    """Returns (y, cache) for inputs x [n x in]."""
    r = bf16 if bf16_mode else (lambda v: np.asarray(v, dtype=np.float64))
    layers = unpack(params, dims)
    a = r(x)
    acts, zs = [a], []
    for i, (W, b) in enumerate(layers):
        z = a @ r(W).T + b
        zs.append(z)
        if i + 1 < len(layers):
            a = r(gelu(z))
            acts.append(a)
    return zs[-1], (acts, zs)


def loss_and_grad(params, dims, x, t, bf16_mode=True):
    """MSE loss and its gradient (flat, parameter order)."""
    r = bf16 if bf16_mode else (lambda v: np.asarray(v, dtype=np.float64))
    layers = unpack(params, dims)
    y, (acts, zs) = forward(params, dims, x, bf16_mode)
    d = y - t
    count = d.size
    loss = float(np.mean(d * d))
    dz = r(2.0 * d / count)
    grads = [None] * len(layers)
    for i in range(len(layers) - 1, -1, -1):
        W, _ = layers[i]
        gW = dz.T @ acts[i]
        gb = dz.sum(0)
        grads[i] = (gW, gb)
        if i > 0:
            dz = r((dz @ r(W)) * gelu_grad(zs[i - 1]))
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat


def adam(params, grads, m, v, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step (step count t >= 1); returns new (params, m, v)."""
    m = beta1 * m + (1.0 - beta1) * grads
    v = beta2 * v + (1.0 - beta2) * grads * grads
    mh = m / (1.0 - beta1 ** t)
    vh = v / (1.0 - beta2 ** t)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v


def init_params(dims, seed):
    """The GPU's initial weights: splitmix64 uniforms, U(-1/sqrt(in), 1/sqrt(in))."""
    mask = (1 << 64) - 1
    s = seed & mask

    def uni():
        nonlocal s
        s = (s + 0x9E3779B97F4A7C15) & mask
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        z ^= z >> 31
        return (z >> 11) * (1.0 / 9007199254740992.0)

    out = []
    for i in range(len(dims) - 1):
        bound = 1.0 / math.sqrt(dims[i])
        W = np.array([[(2.0 * uni() - 1.0) * bound for _ in range(dims[i])] for _ in range(dims[i + 1])])
        b = np.array([(2.0 * uni() - 1.0) * bound for _ in range(dims[i + 1])])
        out.append(np.concatenate([W.ravel(), b]))
    return np.concatenate(out).astype(np.float32)
