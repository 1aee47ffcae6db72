"""GPU checks of the surrogate MLP kernels (BF16 MFMA GEMMs, fused GELU,
loss, bias gradients, Adam) against the numpy oracle, through the C ABI.

Bars: forward outputs within 2e-3 of the BF16-rounding oracle (relative to
the output scale; FP32 vs FP64 sums can move a BF16 rounding by one unit);
gradients within 2e-2 (norm) and cosine > 0.9999 of it; against exact FP64
backpropagation within 5e-2 (the BF16 quantisation itself); Adam to FP32
precision."""
import numpy as np
import pytest

import mlp_oracle as M

pytestmark = pytest.mark.gpu

DIMS = (5, 128, 128, 64)
B = 256


@pytest.fixture(scope="module")
def surr(lcurve_rs):
    from lcurve_rs import surrogate
    return surrogate


def _data(seed, n=B):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, DIMS[0])).astype(np.float32)
    y = (np.sin(np.outer(x @ np.array([1.0, -0.5, 0.3, 0.8, -1.2]), np.linspace(0.5, 3, DIMS[-1])))
         + 0.1 * rng.standard_normal((n, DIMS[-1]))).astype(np.float32)
    return x, y


def test_initial_parameters_match_oracle(surr):
    net = surr.Mlp(DIMS, seed=3, max_batch=B)
    assert net.n_params == len(M.init_params(DIMS, 3))
    assert np.array_equal(net.get_params(), M.init_params(DIMS, 3))


def test_forward_matches_oracle(surr):
    net = surr.Mlp(DIMS, seed=4, max_batch=B)
    x, _ = _data(1, 300)  # not a multiple of 64: predict pads
    y = net.predict(x)
    ref, _ = M.forward(net.get_params().astype(np.float64), DIMS, x, bf16_mode=True)
    scale = np.abs(ref).max()
    assert np.max(np.abs(y - ref)) <= 2e-3 * scale


def test_gradients_match_oracle(surr):
    net = surr.Mlp(DIMS, seed=5, max_batch=B)
    x, t = _data(2)
    loss, g = net.loss_grad(x, t)
    p = net.get_params().astype(np.float64)
    lo, go = M.loss_and_grad(p, DIMS, x, t, bf16_mode=True)
    assert loss == pytest.approx(lo, rel=1e-4)
    assert np.linalg.norm(g - go) <= 2e-2 * np.linalg.norm(go)
    assert np.dot(g, go) / (np.linalg.norm(g) * np.linalg.norm(go)) > 0.9999
    le, ge = M.loss_and_grad(p, DIMS, x, t, bf16_mode=False)
    assert np.linalg.norm(g - ge) <= 5e-2 * np.linalg.norm(ge)
    # bias gradients are sums of the same BF16 values: tight
    o = DIMS[1] * DIMS[0]
    assert np.allclose(g[o:o + DIMS[1]], go[o:o + DIMS[1]], rtol=2e-2, atol=1e-6 * np.abs(go).max())


def test_adam_step_matches_oracle(surr):
    net = surr.Mlp(DIMS, seed=6, max_batch=B)
    x, t = _data(3)
    p0 = net.get_params().astype(np.float64)
    _, g = net.loss_grad(x, t)
    net.adam_step(1e-3)
    p1 = net.get_params()
    ref, _, _ = M.adam(p0, g.astype(np.float64), np.zeros_like(p0), np.zeros_like(p0), 1, 1e-3)
    assert np.allclose(p1, ref, rtol=1e-6, atol=1e-7)


def test_training_reduces_loss(surr):
    net = surr.Mlp(DIMS, seed=7, max_batch=B)
    x, y = _data(4, 4096)
    net.set_data(x, y)
    rng = np.random.default_rng(0)
    first = net.train_epoch(rng.permutation(4096), B, 2e-3)
    for _ in range(30):
        last = net.train_epoch(rng.permutation(4096), B, 2e-3)
    assert last < 0.2 * first
    held_x, held_y = _data(5, 512)
    err = np.mean((net.predict(held_x) - held_y) ** 2)
    assert err < 0.3 * np.mean((held_y - held_y.mean(0)) ** 2)


def test_step_timing_reports_gemm_flops(surr):
    net = surr.Mlp((5, 512, 512, 512, 512, 256), seed=8, max_batch=4096)
    x, y = _data(6, 4096)
    y = np.tile(y, (1, 4))
    net.set_data(x, y)
    t = net.time_step(4096, reps=3)
    assert t["ms_step"] > 0 and t["ms_gemm"] > 0
    # 2*B*(64*512*2 + 3*(512*512*3) + 512*256*3)
    assert t["gemm_flop"] == pytest.approx(2 * 4096 * (64 * 512 * 2 + 3 * 512 * 512 * 3 + 512 * 256 * 3))


def test_full_network_matches_oracle(surr):
    """The configured network, 5-512x4-256 at batch 4096 (the weight
    gradients then split K = 4096 across slices): forward outputs and
    gradients against the BF16-rounding oracle and exact FP64
    backpropagation, at the bars above."""
    dims = (5, 512, 512, 512, 512, 256)
    net = surr.Mlp(dims, seed=9, max_batch=4096)
    rng = np.random.default_rng(7)
    x = rng.uniform(-1, 1, (4096, 5)).astype(np.float32)
    t = (np.sin(np.outer(x @ np.array([1.0, -0.5, 0.3, 0.8, -1.2]), np.linspace(0.5, 3, 256)))
         + 0.1 * rng.standard_normal((4096, 256))).astype(np.float32)
    p = net.get_params().astype(np.float64)
    y = net.predict(x)
    ref, _ = M.forward(p, dims, x, bf16_mode=True)
    assert np.max(np.abs(y - ref)) <= 2e-3 * np.abs(ref).max()
    loss, g = net.loss_grad(x, t)
    lo, go = M.loss_and_grad(p, dims, x, t, bf16_mode=True)
    assert loss == pytest.approx(lo, rel=1e-4)
    assert np.linalg.norm(g - go) <= 2e-2 * np.linalg.norm(go)
    assert np.dot(g, go) / (np.linalg.norm(g) * np.linalg.norm(go)) > 0.9999
    _, ge = M.loss_and_grad(p, dims, x, t, bf16_mode=False)
    assert np.linalg.norm(g - ge) <= 5e-2 * np.linalg.norm(ge)
