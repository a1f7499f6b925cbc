"""Runs of circular SIS halos (e = 0) form their gradient and Hessian terms from the lab-frame offsets
(no rotation to the halo frame and back). A run takes that path only when every halo of it is a
circular SIS -- in a batch, in every sample -- and staging new samples updates the flag with the
records. Each case is checked against the oracle (FP64 criterion of tests/parity.py): a run flagged
circular by mistake would treat an elliptical SIS as round, far outside the tolerance."""
import numpy as np
import pytest

import ltmodel as M
from parity import EXT_SHARE, assert_close, hessian_reference, print_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    import lthip
    lthip.check(lthip.lib().lt_init(0))
    return lthip


def _lens(sis_eps, seed=3):
    rng = np.random.default_rng(seed)
    halos = [dict(type=81, x=0.5, y=-1.0, sigma=900.0, e=0.3, theta=0.7, rcore=6.0, rcut=400.0)]
    for e in sis_eps:
        halos.append(dict(type=5, x=float(rng.uniform(-30, 30)), y=float(rng.uniform(-30, 30)),
                          sigma=float(rng.uniform(120, 300)), e=float(e), theta=float(rng.uniform(0, np.pi))))
    return M.LensModel.from_halos(halos)


def _check_hessian(oracle, lens, grid, a, b, c, name):
    (ea, eb, ec), (ta, tb, tc) = hessian_reference(oracle, lens, grid)
    assert_close(name + " a", a, ea, np.zeros_like(ta), ta)
    assert_close(name + " b", b, eb, np.zeros_like(tb), tb)
    assert_close(name + " c", c, ec, np.zeros_like(tc), tc)


def _batch_hessian(hip, m, grid, nsamples):
    planes = [hip.DeviceArray(nsamples * grid.npix) for _ in range(4)]
    m.map(hip.MAP_HESSIAN, 0, [p.ptr for p in planes], grid)
    hip.sync()
    out = [p.download().reshape(nsamples, grid.ny, grid.nx) for p in planes]
    for p in planes:
        p.close()
    return out


@pytest.mark.parametrize("sis_eps", [(0.0, 0.0, 0.0, 0.0), (0.0, 0.12, 0.0, 0.0), (0.2, 0.3, 0.1, 0.25)])
def test_circular_and_elliptical_runs(hip, oracle, sis_eps):
    lens = _lens(sis_eps)
    grid = M.Grid(M.grid_param(-40, 40, -35, 35), 89, 71)
    print_case(f"SIS runs eps={sis_eps}", lens, grid)
    a, b, c, _ = hip.hessian_grid(lens, grid)
    _check_hessian(oracle, lens, grid, a, b, c, "hessian")
    gx, gy = hip.gradient_grid(lens, grid)
    ex, ey = oracle.gradient_grid(lens, grid, extended=True)
    _, _, ax, ay = oracle.gradient_grid(lens, grid, with_abs=True)
    cg, _ = oracle.cond_planes(lens, grid)
    assert_close("gx", gx, ex, ax, EXT_SHARE * cg)
    assert_close("gy", gy, ey, ay, EXT_SHARE * cg)
    # FP32C fields map: det A within 1e-6 of the FP64 one
    cosmo = M.flat_lcdm(70.0, 0.3)
    f64 = hip.map_fields(lens, grid, cosmo, 0.4, 2.0)
    f32 = hip.map_fields(lens, grid, cosmo, 0.4, 2.0, precision=hip.PREC_FP32C)
    assert np.abs(f32[3] - f64[3]).max() <= 1e-6


def test_batch_flag_needs_every_sample(hip, oracle):
    """Sample 0 has only circular SIS, sample 1 one elliptical SIS: the batch's SIS run is not
    circular, and both samples match the oracle."""
    A = _lens((0.0, 0.0, 0.0, 0.0))
    B = _lens((0.0, 0.0, 0.15, 0.0))
    grid = M.Grid(M.grid_param(-40, 40, -35, 35), 64, 48)
    print_case("SIS batch [circular, elliptical]", A, grid)
    m = hip.Model([A, B], hip.MAP_HESSIAN)
    try:
        a, b, c, _ = _batch_hessian(hip, m, grid, 2)
        for k, L in enumerate((A, B)):
            _check_hessian(oracle, L, grid, a[k], b[k], c[k], f"sample {k}")
    finally:
        m.close()


def test_staging_updates_flag(hip, oracle):
    """A batch created with circular SIS, then staged with an elliptical one, then circular again:
    the maps follow the staged samples (the flag travels with the staged records)."""
    A = _lens((0.0, 0.0, 0.0, 0.0))
    B = _lens((0.0, 0.0, 0.0, 0.3))
    grid = M.Grid(M.grid_param(-40, 40, -35, 35), 64, 48)
    print_case("SIS staging circular -> elliptical -> circular", A, grid)
    m = hip.Model([A, A], hip.MAP_HESSIAN)
    try:
        first = _batch_hessian(hip, m, grid, 2)
        for L, name in ((B, "elliptical"), (A, "circular")):
            m.stage([L, L])
            m.upload()
            a, b, c, _ = _batch_hessian(hip, m, grid, 2)
            for k in range(2):
                _check_hessian(oracle, L, grid, a[k], b[k], c[k], f"{name} sample {k}")
        # back to the circular samples: bit for bit the first maps
        for p, q in zip(first, (a, b, c)):
            np.testing.assert_array_equal(p, q)
    finally:
        m.close()
