"""GPU checks that freed batches give their device memory back.

An ensemble sampler builds, runs and frees one batch per half-step of
exact evaluations, thousands per chain. Each freed batch must leave the
device's free memory where it was. On ROCm 7.2, tables freed right after a
graph launch with only the launching stream synchronised stayed allocated
(about 230 MB of every 128-set batch) until the device ran out; the engine's
release paths now wait on the whole device before freeing.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TRUTH = np.array([0.5, 82.0, 0.015, 15000.0, 3500.0])


def _hip():
    """The HIP runtime the engine already loaded (same soname, same handle)."""
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    return hip


def _free_bytes(hip):
    assert hip.hipDeviceSynchronize() == 0
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(f), ctypes.byref(t)) == 0
    return f.value


def _target(lc):
    from lcurve_rs import mcmc
    times = np.linspace(-0.2, 1.2, 3000)
    return mcmc.Target(mcmc.surrogate_base_text((40, 20, 40, 40)), TRUTH, times, 0.005, seed=42,
                       expose=0.001, ndiv=5)


def test_freed_batches_return_their_memory(lcurve_rs):
    lc = lcurve_rs
    from lcurve_rs import mcmc
    tg = _target(lc)
    hip = _hip()
    rng = np.random.default_rng(3)

    def one(measure=False):
        p = TRUTH * (1.0 + 1e-4 * rng.standard_normal((48, 5)))
        before = _free_bytes(hip) if measure else 0
        b = lc.Batch(tg.model, tg.times, tg.expose, tg.ndiv, names=mcmc.PARAM_NAMES, values=p, flux=tg.y,
                     ferr=tg.ferr, weight=tg.w)
        size = before - _free_bytes(hip) if measure else 0
        chi = b.run(scale=True, want_flux=False)["chisq"]
        del b  # no synchronisation between the run and the free, as in Target.logp_exact
        assert np.all(np.isfinite(chi))
        return size

    size = one(measure=True)  # also takes the runtime's one-time allocations
    assert size > 32 << 20, size
    f0 = _free_bytes(hip)
    for _ in range(6):
        one()
    lost = f0 - _free_bytes(hip)
    print(f"batch of 48 sets: {size / 2**20:.0f} MB; lost over 6 freed batches: {lost / 2**20:.0f} MB")
    assert lost < size // 2, (lost, size)


def test_freed_device_packs_return_their_memory(lcurve_rs):
    lc = lcurve_rs
    from lcurve_rs import mcmc
    tg = _target(lc)
    hip = _hip()
    rng = np.random.default_rng(4)

    def one(measure=False):
        models = []
        for row in TRUTH * (1.0 + 1e-4 * rng.standard_normal((16, 5))):
            m = lc.Model.from_text(tg.base_text)
            for name, v in zip(mcmc.PARAM_NAMES, row):
                m.set_param(name, v)
            models.append(m)
        before = _free_bytes(hip) if measure else 0
        dp = lc.DevicePackBatch(models, tg.times, tg.expose, tg.ndiv, flux=tg.y, ferr=tg.ferr, weight=tg.w)
        out = dp.run(scale=True, want_flux=False)
        size = before - _free_bytes(hip) if measure else 0
        del dp
        assert np.all(np.isfinite(out["chisq"]))
        return size

    size = one(measure=True)
    assert size > 8 << 20, size
    f0 = _free_bytes(hip)
    for _ in range(4):
        one()
    lost = f0 - _free_bytes(hip)
    print(f"device pack of 16 sets: {size / 2**20:.0f} MB; lost over 4 freed packs: {lost / 2**20:.0f} MB")
    assert lost < size // 2, (lost, size)
