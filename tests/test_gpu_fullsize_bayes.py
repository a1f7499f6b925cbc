"""The batched Bayes configuration at its full size: 256 perturbed samples of the 1 dPIE + 10 SIS
cluster lens, each a 2048 x 2048 amplification map (mu, mode 1), in one launch.

Checked through properties that do not need the oracle at this size: sampled members of the batch
equal the single-model map of the same sample bit for bit (FP64 and FP32C: the same per-pixel code
runs whatever the batch), det A = 1/mu is finite at every pixel, and the FP32C batch's det A is
within 1e-6 of the FP64 batch's at every pixel of every sample (the bound the FP32-compensated maps
are held to)."""
import numpy as np
import pytest

import ltmodel as M

pytestmark = pytest.mark.gpu

S, N = 256, 2048


@pytest.fixture(scope="module")
def hip():
    import lthip
    lthip.check(lthip.lib().lt_init(0))
    return lthip


def test_bayes_batch_fullsize(hip):
    base = M.group_lens(seed=1)
    samples = M.perturbed_samples(base, S, seed=100)
    grid = M.Grid(M.grid_param(-100.0, 100.0, -100.0, 100.0), N, N)
    r = hip.cosmo_dlsds(0.4, 2.0, M.flat_lcdm(70.0, 0.3))
    print(f"[case] Bayes batch full size: {S} samples x {N}x{N}, {base.n} halos", flush=True)
    s = hip.stream_create()
    model = hip.Model(samples, hip.MAP_HESSIAN)
    mu = {p: hip.DeviceArray(S * grid.npix) for p in (0, 1)}
    try:
        for p in (0, 1):
            model.map(hip.MAP_AMPLI, 1, [mu[p].ptr], grid, r=r, precision=p, stream=s)
        hip.sync(s)
        worst = 0.0
        for k in range(S):
            d64 = 1.0 / mu[0].download_range(k * grid.npix, grid.npix)
            d32 = 1.0 / mu[1].download_range(k * grid.npix, grid.npix)
            assert np.isfinite(d64).all() and np.isfinite(d32).all()
            worst = max(worst, float(np.abs(d32 - d64).max()))
        print(f"max |det A (FP32C) - det A (FP64)| over {S} samples: {worst:.3e}", flush=True)
        assert worst <= 1e-6
        one = hip.DeviceArray(grid.npix)
        try:
            for k in (0, 101, S - 1):
                single = hip.Model(samples[k], hip.MAP_HESSIAN)
                for p in (0, 1):
                    single.map(hip.MAP_AMPLI, 1, [one.ptr], grid, r=r, precision=p, stream=s)
                    hip.sync(s)
                    np.testing.assert_array_equal(mu[p].download_range(k * grid.npix, grid.npix), one.download())
                single.close()
        finally:
            one.close()
    finally:
        for a in mu.values():
            a.close()
        model.close()
        hip.stream_destroy(s)
