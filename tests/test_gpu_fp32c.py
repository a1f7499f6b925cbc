"""FP32-compensated Hessian maps (LT_PREC_FP32C) against FP64.

The bound is the one the FP32 path is built for: max |delta det A| <= 1e-6 in absolute terms
(det A = (1 - kappa)^2 - gamma^2 with kappa, gamma scaled by dlsds), against the FP64 path, which
itself equals the oracle to 1e-12 relative (test_gpu_parity.py). Hessian components are held to
an absolute 2e-7 (their scale in these lenses is 0.1-10 near cores)."""
import numpy as np
import pytest

import ltmodel as M
from parity import print_case

pytestmark = pytest.mark.gpu

DET_TOL = 1e-6
HESS_TOL = 2e-7


@pytest.fixture(scope="module")
def hip():
    import lthip
    lthip.check(lthip.lib().lt_init(0))
    return lthip


def _mixed_lens():
    rng = np.random.default_rng(31)
    halos = [dict(type=81, x=0.0, y=0.0, sigma=900.0, e=0.4, theta=0.3, rcore=8.0, rcut=900.0)]
    for _ in range(40):
        m = rng.uniform(18, 24)
        f = 10 ** (0.4 * (18 - m))
        halos.append(dict(type=81, x=float(rng.uniform(-60, 60)), y=float(rng.uniform(-60, 60)),
                          sigma=180.0 * f ** 0.25, e=float(rng.uniform(0, 0.6)), theta=float(rng.uniform(0, 3.1)),
                          rcore=0.15 * f ** 0.5, rcut=40.0 * f ** 0.5))
    for _ in range(6):
        halos.append(dict(type=8, x=float(rng.uniform(-60, 60)), y=float(rng.uniform(-60, 60)),
                          sigma=float(rng.uniform(100, 250)), e=float(rng.uniform(0.05, 0.5)),
                          theta=float(rng.uniform(0, 3.1)), rcore=float(rng.uniform(0.05, 0.5)), rcut=0.0))
    for _ in range(8):
        halos.append(dict(type=5, x=float(rng.uniform(-60, 60)), y=float(rng.uniform(-60, 60)),
                          sigma=float(rng.uniform(80, 250)), e=0.0, theta=float(rng.uniform(0, 3.1)),
                          rcore=0.0, rcut=0.0))
    halos.append(dict(type=14, x=0.0, y=0.0, b0=0.0, epot=0.05, theta=0.7, rcore=0.0, rcut=0.0))
    return M.LensModel.from_halos(sorted(halos, key=lambda h: h["type"]))


def _detA(a, b, c, r):
    k = 0.5 * (a + c) * r
    g1 = 0.5 * (a - c) * r
    g2 = b * r
    return (1 - k) ** 2 - (g1 * g1 + g2 * g2)


def test_fp32c_mixed_types_vs_oracle(hip, oracle):
    """All halo types (SIS, PIEMD, dPIE, external shear) at every pixel of a 256^2 grid over the
    galaxies, near and far from their cores, against the extended-precision oracle."""
    lens = _mixed_lens()
    grid = M.Grid(M.grid_param(-64.0, 64.0, -64.0, 64.0), 256, 256)
    print_case("fp32c mixed types", lens, grid)
    a, b, c, d = hip.hessian_grid(lens, grid, precision=hip.PREC_FP32C)
    ea, eb, ec, ed = oracle.hessian_planes(lens, grid, extended=True)
    assert np.abs(a - ea).max() <= HESS_TOL
    assert np.abs(b - eb).max() <= HESS_TOL
    assert np.abs(c - ec).max() <= HESS_TOL
    np.testing.assert_array_equal(b, d)
    r = 0.7
    assert np.abs(_detA(a, b, c, r) - _detA(ea, eb, ec, r)).max() <= DET_TOL


def test_fp32c_near_removable_singularity(hip, oracle):
    """Pixels within 0.5" of the halo-frame point (0, 2 rcut sqrt(e)) of a galaxy, where the
    reference's closed form is 0/0, taken by the FP32 form (the point lies outside the galaxy's FP64
    radius): no loss of accuracy there. (The grid has an even size so that no pixel sits exactly on
    the point, where the oracle's closed form itself is 0/0.)"""
    lens = M.LensModel.from_halos([dict(type=81, x=3.0, y=-2.0, sigma=100.0, e=0.5, theta=0.9, rcore=0.1,
                                        rcut=30.0)])
    eps = lens.epot[0]
    ys = 2 * 30.0 * np.sqrt(eps)
    assert ys > 20.0 * lens.b0[0] * 30.0 / 29.9  # outside the FP64 radius
    th = lens.angle[0]
    cx = 3.0 - ys * np.sin(th)
    cy = -2.0 + ys * np.cos(th)
    grid = M.Grid(M.grid_param(cx - 0.5, cx + 0.5, cy - 0.5, cy + 0.5), 96, 96)
    print_case("fp32c removable singularity", lens, grid)
    a, b, c, _ = hip.hessian_grid(lens, grid, precision=hip.PREC_FP32C)
    ea, eb, ec, _ = oracle.hessian_planes(lens, grid, extended=True)
    for got, ex in ((a, ea), (b, eb), (c, ec)):
        assert np.abs(got - ex).max() <= 2e-8  # the closed form in FP32 is off by ~1 here


def test_fp32c_kmap_and_map_kinds(hip, oracle):
    lens = _mixed_lens()
    lens.dlsds[:] = np.where(np.arange(lens.n) % 3 == 0, 0.6, 0.75)
    lens._soa = None
    grid = M.Grid(M.grid_param(-50.0, 50.0, -50.0, 50.0), 128, 96)
    k32 = hip.hessian_grid(lens, grid, precision=hip.PREC_FP32C, kmap=True)
    k64 = hip.hessian_grid(lens, grid, kmap=True)
    for p32, p64 in zip(k32, k64):
        assert np.abs(p32 - p64).max() <= HESS_TOL
    cosmo = M.flat_lcdm()
    for kind, modes in ((hip.MAP_AMPLI, (1, 2, 5, 6)), (hip.MAP_SHEAR, (1, 3, 4)), (hip.MAP_MASS, (1, 2))):
        for mode in modes:
            m32 = hip.map_grid(kind, mode, lens, grid, cosmo, 0.4, 2.0, precision=hip.PREC_FP32C)
            m64 = hip.map_grid(kind, mode, lens, grid, cosmo, 0.4, 2.0)
            assert np.isfinite(m32).all()
            if kind == hip.MAP_AMPLI and mode in (1, 2):
                # mu = 1/det A: compare det A
                assert np.abs(1 / m32 - 1 / m64).max() <= DET_TOL
            else:
                assert np.abs(m32 - m64).max() <= 1e-6 * max(1.0, np.abs(m64).max() / 1e3)
    # displacement maps in FP32C come from the FP32C gradient kernel (bound of
    # tests/test_gpu_fp32c_gradient.py, 2e-5 arcsec on the deflection; no type-14 gradient)
    glens = lens.subset(np.nonzero(lens.type != 14)[0])
    g32 = hip.map_grid(hip.MAP_DPL, 1, glens, grid, cosmo, 0.4, 2.0, precision=hip.PREC_FP32C)
    g64 = hip.map_grid(hip.MAP_DPL, 1, glens, grid, cosmo, 0.4, 2.0)
    assert np.hypot(g32[0] - g64[0], g32[1] - g64[1]).max() <= 2e-5


def test_fp32c_cluster_full_size(hip):
    """The cluster config at full size (4096^2, 503 dPIE halos): det A of the FP32C fields map within
    1e-6 of the FP64 map at every pixel; kappa and gamma within 2e-7."""
    lens = M.cluster_lens()
    grid = M.cluster_grid(4096)
    print_case("fp32c cluster lens full size", lens, grid)
    cosmo = M.flat_lcdm()
    f64 = hip.map_fields(lens, grid, cosmo, 0.4, 2.0)
    f32 = hip.map_fields(lens, grid, cosmo, 0.4, 2.0, precision=hip.PREC_FP32C)
    err = {}
    for name, p32, p64 in zip(("kappa", "gamma1", "gamma2", "detA"), f32[:4], f64[:4]):
        err[name] = float(np.abs(p32 - p64).max())
    print("[fp32c cluster lens] max abs errors:", err, flush=True)
    assert err["detA"] <= DET_TOL
    for k in ("kappa", "gamma1", "gamma2"):
        assert err[k] <= HESS_TOL
