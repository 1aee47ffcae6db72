"""CPU checks of the surrogate MLP oracle (oracle/mlp_oracle.py): BF16
rounding, GELU and its derivative, backpropagation against central finite
differences in FP64, and Adam's first step in closed form."""
import math

import numpy as np
import pytest

import mlp_oracle as M

DIMS = (5, 16, 16, 8)


def test_bf16_round_to_nearest_even():
    assert M.bf16(1.0) == 1.0
    assert M.bf16(1.0 + 2.0 ** -8) == 1.0            # halfway, ties to even
    assert M.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6   # halfway, rounds up to even
    assert M.bf16(1.0 + 3 * 2.0 ** -9) == 1.0 + 2.0 ** -7
    assert M.bf16(-2.5) == -2.5
    x = np.random.default_rng(0).standard_normal(1000)
    r = M.bf16(x)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8 + 1e-300)


def test_gelu_values_and_derivative():
    assert M.gelu(np.array([0.0]))[0] == 0.0
    assert M.gelu(np.array([1.0]))[0] == pytest.approx(0.8413447460685429, rel=1e-15)
    z = np.linspace(-5, 5, 41)
    h = 1e-6
    fd = (M.gelu(z + h) - M.gelu(z - h)) / (2 * h)
    assert np.allclose(M.gelu_grad(z), fd, rtol=1e-8, atol=1e-9)


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = M.init_params(DIMS, 5).astype(np.float64)
    x = rng.uniform(-1, 1, (12, DIMS[0]))
    t = rng.standard_normal((12, DIMS[-1]))
    _, g = M.loss_and_grad(p, DIMS, x, t, bf16_mode=False)
    idx = rng.choice(len(p), 40, replace=False)
    for i in idx:
        e = np.zeros_like(p)
        e[i] = 1e-6
        lp, _ = M.loss_and_grad(p + e, DIMS, x, t, bf16_mode=False)
        lm, _ = M.loss_and_grad(p - e, DIMS, x, t, bf16_mode=False)
        fd = (lp - lm) / 2e-6
        assert g[i] == pytest.approx(fd, rel=1e-6, abs=1e-10), i


def test_adam_first_step_is_signed_lr():
    g = np.array([0.3, -2.0, 1e-3])
    p, m, v = M.adam(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, lr=0.01)
    assert np.allclose(p, -0.01 * g / (np.abs(g) + 1e-8), rtol=1e-12)


def test_init_params_bounds_and_repeatability():
    p1 = M.init_params(DIMS, 9)
    p2 = M.init_params(DIMS, 9)
    assert np.array_equal(p1, p2)
    o = 0
    for i in range(len(DIMS) - 1):
        n = DIMS[i + 1] * DIMS[i] + DIMS[i + 1]
        assert np.all(np.abs(p1[o:o + n]) <= 1.0 / math.sqrt(DIMS[i]))
        o += n
    assert o == len(p1)
