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
# FP32C potential: |phi(FP32C) - phi(FP64)| <= POT_REL sum_h |phi_h| at every pixel
POT_REL = 1e-5


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


def _pot_map(hip, lens, grid, precision, px=0):
    m = hip.Model(lens, hip.MAP_POT)
    out = hip.DeviceArray(grid.npix)
    m.map(hip.MAP_POT, 3, [out.ptr], grid, precision=precision, px=px)
    hip.sync()
    v = out.download().reshape(grid.ny, grid.nx)
    m.close()
    out.close()
    return v


@pytest.mark.parametrize("which", ["mixed", "cluster", "centre"])
def test_fp32c_potential(hip, oracle, which):
    """The FP32C potential against the FP64 kernel (itself checked against the extended oracle in
    test_gpu_parity.py): the error, relative to the sum over halos of |phi_h| at each pixel, stays
    below POT_REL; bit-identical for 1, 2 and 4 pixels per thread. "centre": pixels 1e-3 to 3e-2
    core radii from a dPIE centre, where the first logarithm is a log1p of a small quantity."""
    if which == "mixed":
        lens = _mixed_lens()
        lens = lens.subset(np.nonzero(lens.type != 14)[0])  # external shear has no potential
        grid = M.Grid(M.grid_param(-70, 70, -65, 65), 211, 199)
    elif which == "centre":
        lens = M.LensModel.from_halos(
            [dict(type=81, x=0.37, y=-0.21, sigma=420.0, e=0.35, theta=0.6, rcore=0.5, rcut=60.0),
             dict(type=8, x=8.0, y=5.0, sigma=250.0, e=0.2, theta=1.9, rcore=0.2, rcut=30.0)])
        n, step = 33, 1e-3
        x0, y0 = 0.37 - (n // 2 + 0.5) * step, -0.21 - (n // 2 + 0.5) * step
        grid = M.Grid(M.grid_param(x0, x0 + (n - 1) * step, y0, y0 + (n - 1) * step), n, n)
    else:
        lens = M.cluster_lens(seed=2)
        grid = M.Grid(M.grid_param(-150, 150, -150, 150), 160, 150)
    print_case(f"FP32C potential {which}", lens, grid)
    f64 = _pot_map(hip, lens, grid, hip.PREC_FP64)
    f32 = _pot_map(hip, lens, grid, hip.PREC_FP32C)
    S = np.zeros_like(f64)
    for i in range(lens.n):
        S += np.abs(oracle.potential_grid(lens.subset([i]), grid))
    # the I_0.5 potential is 0/0 at a halo centre (the reference's gives NaN there too): both kernels
    # give NaN at exactly the same pixels, and are compared everywhere else
    fin = np.isfinite(f64)
    np.testing.assert_array_equal(np.isfinite(f32), fin)
    rel = np.abs(f32 - f64)[fin] / S[fin]
    print(f"[fp32c potential {which}] {(~fin).sum()} centre pixels; max |d phi| / sum |phi_h| = {rel.max():.3e}, "
          f"median {np.median(rel):.3e}, max |d phi| = {np.abs(f32 - f64)[fin].max():.3e} of max |phi| "
          f"{np.abs(f64[fin]).max():.3e}", flush=True)
    assert rel.max() <= POT_REL
    for px in (1, 4):
        np.testing.assert_array_equal(_pot_map(hip, lens, grid, hip.PREC_FP32C, px=px), f32)


def _fields_px(hip, model, grid, px, nsamples=1, kind=None):
    kind = hip.MAP_FIELDS if kind is None else kind
    outs = [hip.DeviceArray(nsamples * grid.npix) for _ in range(5)]
    model.map(kind, 0, [o.ptr for o in outs], grid, r=0.7, precision=hip.PREC_FP32C, px=px)
    hip.sync()
    v = [o.download().reshape(nsamples, grid.ny, grid.nx) for o in outs]
    for o in outs:
        o.close()
    return v


def test_fp32c_halo_lane_shape(hip, oracle):
    """The lanes-over-halos launch shape of the FP32C Hessian (LT_PX_HALO_LANES): every halo type and a
    run longer than 64 halos, grid sizes that are not multiples of the 16-pixel segment or the 4-row
    work-group, a batch of samples; det A within DET_TOL and the Hessian within HESS_TOL of the
    extended oracle, as the pixel-per-lane shape."""
    base = _mixed_lens()
    extra = M.cluster_lens(seed=8, ngal=90)          # a 93-halo dPIE run: two chunks of 64 lanes
    lens = M.LensModel.from_halos(
        [dict(type=int(t), x=float(x), y=float(y), b0=float(b), epot=float(e), theta=float(a), rcore=float(rc),
              rcut=float(rt)) for L in (base, extra)
         for t, x, y, b, e, a, rc, rt in zip(L.type, L.x, L.y, L.b0, L.epot, L.angle, L.rcore, L.rcut)])
    grid = M.Grid(M.grid_param(-70.0, 70.0, -65.0, 65.0), 203, 129)
    print_case("fp32c halo-lane shape", lens, grid)
    m = hip.Model(lens, hip.MAP_HESSIAN)
    hl = _fields_px(hip, m, grid, hip.PX_HALO_LANES, kind=hip.MAP_HESSIAN)
    m.close()
    ea, eb, ec, _ = oracle.hessian_planes(lens, grid, extended=True)
    for got, ex in ((hl[0][0], ea), (hl[1][0], eb), (hl[2][0], ec)):
        assert np.abs(got - ex).max() <= HESS_TOL
    np.testing.assert_array_equal(hl[1], hl[3])
    r = 0.7
    assert np.abs(_detA(hl[0][0], hl[1][0], hl[2][0], r) - _detA(ea, eb, ec, r)).max() <= DET_TOL
    # a batch of samples: each sample's map equals the single-model launch of the same shape
    samples = M.perturbed_samples(M.group_lens(seed=4), 3, seed=13)
    g2 = M.Grid(M.grid_param(-100.0, 100.0, -100.0, 100.0), 97, 61)
    bm = hip.Model(samples, hip.MAP_HESSIAN)
    batch = _fields_px(hip, bm, g2, hip.PX_HALO_LANES, nsamples=3)
    bm.close()
    for k, L in enumerate(samples):
        sm = hip.Model(L, hip.MAP_HESSIAN)
        one = _fields_px(hip, sm, g2, hip.PX_HALO_LANES)
        sm.close()
        for p in range(5):
            np.testing.assert_array_equal(batch[p][k], one[p][0])
        a, b, c, _ = hip.hessian_grid(L, g2)
        assert np.abs(batch[3][k] - _detA(a, b, c, 0.7)).max() <= DET_TOL
