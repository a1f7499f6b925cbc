"""FP32-compensated gradient and displacement maps (k_gradient32).

Bound: |alpha(FP32C) - alpha(FP64)| <= 2e-5 arcsec at every pixel (the deflection, not a relative
error: deflections reach ~100 arcsec, and 2e-5 arcsec is 0.02 milliarcsecond). The FP64 maps are held
to the oracle by tests/test_gpu_parity.py. Cases: the 503-halo cluster lens, a mixed lens (elliptical
SIS, PIEMD, dPIE), a dPIE around the removable point (0, 2 rcut sqrt(e)) of its cut radius -- where
the reference's N/D form cancels and the FP32 kernel's Q/S form does not -- batch = single, and the
displacement map = D_ls/D_s x gradient."""
import math

import numpy as np
import pytest

import ltmodel as M
from parity import print_case

pytestmark = pytest.mark.gpu

BOUND = 2e-5


@pytest.fixture(scope="module")
def hip():
    import lthip
    lthip.check(lthip.lib().lt_init(0))
    return lthip


def _max_err(hip, lens, grid):
    gx, gy = hip.gradient_grid(lens, grid)
    fx, fy = hip.gradient_grid(lens, grid, precision=hip.PREC_FP32C)
    assert np.isfinite(fx).all() and np.isfinite(fy).all()
    return float(np.hypot(fx - gx, fy - gy).max())


def test_cluster_lens(hip):
    lens = M.cluster_lens(seed=2)
    grid = M.cluster_grid(1024)
    print_case("FP32C gradient, cluster lens", lens, grid)
    err = _max_err(hip, lens, grid)
    print(f"max |dalpha| = {err:.3e} arcsec", flush=True)
    assert err <= BOUND


def test_mixed_lens(hip):
    from test_gpu_parity import mixed_lens
    lens = mixed_lens()
    grid = M.Grid(M.grid_param(-40, 40, -35, 35), 301, 237)
    print_case("FP32C gradient, mixed lens", lens, grid)
    err = _max_err(hip, lens, grid)
    print(f"max |dalpha| = {err:.3e} arcsec", flush=True)
    assert err <= BOUND


def test_cut_radius_removable_point(hip):
    """A fine grid around the halo-frame point (0, 2 rcut sqrt(e)) of a dPIE, rotated to the lab frame.
    With b0 = 0.5 the point (33" out) lies beyond the halo's FP64 near radius (20 b0 rcut / (rcut -
    rcore) = 10"), so every pixel takes the FP32 term. The grid is offset so that no pixel comes closer
    than 1.5e-3" to the point: there the N/D form (the reference's, followed by the FP64 kernel) is
    still accurate in FP64, while in FP32 it would be off by ~0.05"; exactly at the point it is 0/0
    and only the Q/S form has a value."""
    e_pot, theta, rcut = 0.3, 0.6, 30.0
    lens = M.LensModel.from_halos([{"type": 81, "x": 1.0, "y": -2.0, "b0": 0.5, "epot": e_pot, "theta": theta,
                                    "rcore": 0.5, "rcut": rcut}])
    yh = 2 * rcut * math.sqrt(e_pot)
    cx, cy = 1.0 - yh * math.sin(theta), -2.0 + yh * math.cos(theta)  # lab position of (0, yh)
    cx, cy = cx + 0.00123, cy - 0.00087
    grid = M.Grid(M.grid_param(cx - 0.5, cx + 0.5, cy - 0.5, cy + 0.5), 201, 201)
    print_case("FP32C gradient at the cut radius' removable point", lens, grid)
    err = _max_err(hip, lens, grid)
    print(f"max |dalpha| = {err:.3e} arcsec", flush=True)
    assert 0 < err <= BOUND


def test_batch_and_displacement(hip):
    base = M.group_lens(seed=4)
    samples = M.perturbed_samples(base, 3, seed=13)
    grid = M.group_grid(160)
    print_case("FP32C gradient batch of 3 and displacement", base, grid)
    s = hip.stream_create()
    try:
        bm = hip.Model(samples, hip.MAP_GRADIENT)
        out = [hip.DeviceArray(3 * grid.npix) for _ in range(2)]
        bm.map(hip.MAP_GRADIENT, 0, [o.ptr for o in out], grid, precision=hip.PREC_FP32C, stream=s)
        hip.sync(s)
        got = [o.download().reshape(3, grid.ny, grid.nx) for o in out]
        dpl = [hip.DeviceArray(3 * grid.npix) for _ in range(2)]
        bm.map(hip.MAP_DPL, 1, [o.ptr for o in dpl], grid, r=0.7, precision=hip.PREC_FP32C, stream=s)
        hip.sync(s)
        dgot = [o.download().reshape(3, grid.ny, grid.nx) for o in dpl]
        for k, L in enumerate(samples):
            fx, fy = hip.gradient_grid(L, grid, precision=hip.PREC_FP32C)
            np.testing.assert_array_equal(got[0][k], fx)
            np.testing.assert_array_equal(got[1][k], fy)
            np.testing.assert_array_equal(dgot[0][k], fx * 0.7)
            np.testing.assert_array_equal(dgot[1][k], fy * 0.7)
        for o in out + dpl:
            o.close()
        bm.close()
    finally:
        hip.stream_destroy(s)



def _tile_lens(n_dpie, n_sis, sis_circular, n_piemd, seed):
    rng = np.random.default_rng(seed)
    halos = []
    for _ in range(n_dpie):
        halos.append(dict(type=81, x=float(rng.uniform(-60, 60)), y=float(rng.uniform(-60, 60)),
                          sigma=float(rng.uniform(80, 250)), e=float(rng.uniform(0, 0.5)),
                          theta=float(rng.uniform(0, math.pi)), rcore=float(rng.uniform(0.1, 2.0)),
                          rcut=float(rng.uniform(10, 80))))
    for _ in range(n_sis):
        halos.append(dict(type=5, x=float(rng.uniform(-60, 60)), y=float(rng.uniform(-60, 60)),
                          sigma=float(rng.uniform(60, 200)),
                          e=0.0 if sis_circular else float(rng.uniform(0.05, 0.4)),
                          theta=float(rng.uniform(0, math.pi))))
    for _ in range(n_piemd):
        halos.append(dict(type=8, x=float(rng.uniform(-60, 60)), y=float(rng.uniform(-60, 60)),
                          sigma=float(rng.uniform(80, 250)), e=float(rng.uniform(0, 0.5)),
                          theta=float(rng.uniform(0, math.pi)), rcore=float(rng.uniform(0.1, 2.0))))
    return M.LensModel.from_halos(halos)


@pytest.mark.parametrize("n_dpie, n_sis, circular, n_piemd", [(64, 65, False, 7), (63, 129, True, 0)])
def test_tile_boundaries(hip, n_dpie, n_sis, circular, n_piemd):
    """The FP32 records of a run are staged through LDS 64 at a time. Runs (one per type) of 64 dPIE
    (one full tile), 65 elliptical SIS (a full tile and one record) and 7 PIEMD (a partial tile), then
    63 dPIE and a circular-SIS run of 129 (two full tiles and one record), on a 203 x 131 grid (131 rows:
    the last work-group has rows past the grid, which stage records and store nothing)."""
    lens = _tile_lens(n_dpie, n_sis, circular, n_piemd, seed=17 + n_sis)
    grid = M.Grid(M.grid_param(-70, 70, -45, 45), 203, 131)
    print_case(f"FP32C gradient across LDS tile boundaries: {n_dpie} dPIE, {n_sis} SIS"
               f" ({'circular' if circular else 'elliptical'}), {n_piemd} PIEMD", lens, grid)
    err = _max_err(hip, lens, grid)
    print(f"max |dalpha| = {err:.3e} arcsec", flush=True)
    assert err <= BOUND
