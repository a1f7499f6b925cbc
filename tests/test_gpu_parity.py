"""GPU parity of the FP64 HIP kernels against the oracle, through the C-ABI.

Criterion (tests/parity.py): |gpu - exact| <= 1e-12 * sum|terms| + FP64 rounding scale of the closed
form, where `exact` is the oracle evaluated in 80-bit extended precision.
"""
import json
import math
import os

import numpy as np
import pytest

import ltmodel as M
from conftest import GOLDEN
from parity import EXT_SHARE, REL, assert_close, hessian_reference, print_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    import lthip
    lthip.check(lthip.lib().lt_init(0))
    return lthip


def mixed_lens(seed=7, n81=5, n8=3, n5=4, n14=0):
    rng = np.random.default_rng(seed)
    halos = []
    for t, k in ((81, n81), (8, n8), (5, n5), (14, n14)):
        for _ in range(k):
            halos.append(dict(type=t, x=float(rng.uniform(-30, 30)), y=float(rng.uniform(-30, 30)),
                              sigma=float(rng.uniform(150, 700)), e=float(rng.uniform(0.05, 0.6)),
                              theta=float(rng.uniform(0, math.pi)), rcore=float(10 ** rng.uniform(-1, 0.7)),
                              rcut=float(10 ** rng.uniform(1, 2.5))))
    if n14:
        for h in halos:
            if h["type"] == 14:
                h["e"] = float(rng.uniform(0.01, 0.1))
                h["epot"] = h["e"]
    return M.LensModel.from_halos(halos)


# ------------------------------------------------------------------------------------------------
# gradient
# ------------------------------------------------------------------------------------------------

def test_gradient_kat1(hip):
    """KAT-1 (Benchmarks/GradientBenchmark/setup.cpp:19-44) as a one-pixel grid."""
    d = json.load(open(os.path.join(GOLDEN, "kat1_gradient.json")))
    n = len(d["type"])
    L = M.LensModel(type=d["type"], x=d["pos_x"], y=d["pos_y"], b0=d["b0"], angle=d["theta"], epot=d["epot"],
                    rcore=d["rcore"], rcut=d["rcut"], z=np.zeros(n), dlsds=np.ones(n))
    grid = M.Grid(M.grid_param(d["x"], d["x"] + 1, d["y"], d["y"] + 1), 1, 1, step=(1.0, 1.0))
    gx, gy = hip.gradient_grid(L, grid)
    assert abs(gx[0, 0] - d["grad_x"]) <= 1e-13 * abs(d["grad_x"])
    assert abs(gy[0, 0] - d["grad_y"]) <= 1e-13 * abs(d["grad_y"])


@pytest.mark.parametrize("istart,jstart", [(0, 0), (5, 3)])
def test_gradient_mixed(hip, oracle, istart, jstart):
    lens = mixed_lens()
    grid = M.Grid(M.grid_param(-40, 40, -35, 35), 97, 83)
    print_case("gradient mixed", lens, grid)
    gx, gy = hip.gradient_grid(lens, grid, istart, jstart)
    ex, ey = oracle.gradient_grid(lens, grid, istart, jstart, extended=True)
    _, _, ax, ay = oracle.gradient_grid(lens, grid, istart, jstart, with_abs=True)
    cg, _ = oracle.cond_planes(lens, grid, istart, jstart)
    assert_close("gx", gx, ex, ax, EXT_SHARE * cg)
    assert_close("gy", gy, ey, ay, EXT_SHARE * cg)


def test_gradient_group_lens(hip, oracle):
    """The group-lens gradient workload at its full size: the whole 512^2 plane (1 dPIE + 10 SIS over
    200"), every pixel against the oracle."""
    lens = M.group_lens(seed=1)
    grid = M.group_grid(512)
    print_case("gradient group lens", lens, grid)
    gx, gy = hip.gradient_grid(lens, grid)
    ex, ey = oracle.gradient_grid(lens, grid, extended=True)
    _, _, ax, ay = oracle.gradient_grid(lens, grid, with_abs=True)
    cg, _ = oracle.cond_planes(lens, grid)
    assert_close("gx", gx, ex, ax, EXT_SHARE * cg)
    assert_close("gy", gy, ey, ay, EXT_SHARE * cg)
    # against the double oracle (the reference's precision: its own rounding, cond) as well
    dx, dy = oracle.gradient_grid(lens, grid)
    assert_close("gx-double", gx, dx, ax, 2 * cg)


# ------------------------------------------------------------------------------------------------
# Hessian
# ------------------------------------------------------------------------------------------------

def test_hessian_reference_golden(hip, oracle):
    """Against the reference's own compiled gradient2_grid_CPU output (tests/golden/hess81_grid.npz)."""
    z = np.load(os.path.join(GOLDEN, "hess81_grid.npz"))
    n = len(z["type"])
    L = M.LensModel(type=z["type"], x=z["x"], y=z["y"], b0=z["b0"], angle=z["angle"], epot=z["epot"],
                    rcore=z["rcore"], rcut=z["rcut"], z=np.full(n, 0.4), dlsds=np.ones(n))
    nx, ny, i0, j0 = (int(v) for v in z["shape"])
    grid = M.Grid(M.grid_param(*z["frame"]), nx, ny, step=tuple(z["dxdy"]))
    a, b, c, d = hip.hessian_grid(L, grid, i0, j0)
    (ea, eb, ec), (ta, tb, tc) = hessian_reference(oracle, L, grid, i0, j0)
    assert_close("a vs exact", a, ea, np.zeros_like(ta), ta)
    assert_close("b vs exact", b, eb, np.zeros_like(tb), tb)
    assert_close("c vs exact", c, ec, np.zeros_like(tc), tc)
    # and against the reference's own FP64 numbers: the reference's closed form carries its full FP64
    # rounding `cond` next to the removable point, the GPU's tolerance on top of it
    _, ch = oracle.cond_planes(L, grid, i0, j0)
    assert_close("a vs reference", a, z["a"], np.zeros_like(ta), ta + REL * np.abs(ea) + ch)
    assert_close("b vs reference", b, z["b"], np.zeros_like(tb), tb + REL * np.abs(eb) + ch)
    assert_close("c vs reference", c, z["c"], np.zeros_like(tc), tc + REL * np.abs(ec) + ch)
    np.testing.assert_array_equal(b, d)


@pytest.mark.parametrize("kmap", [False, True])
def test_hessian_mixed(hip, oracle, kmap):
    lens = mixed_lens(n14=2)
    if kmap:
        lens.dlsds[:] = np.where(lens.type == 81, 0.7, np.where(lens.type == 5, 0.55, 0.62))
        lens._soa = None
    grid = M.Grid(M.grid_param(-40, 40, -35, 35), 101, 77)
    print_case(f"hessian mixed kmap={kmap}", lens, grid)
    a, b, c, d = hip.hessian_grid(lens, grid, 2, 7, kmap=kmap)
    (ea, eb, ec), (ta, tb, tc) = hessian_reference(oracle, lens, grid, 2, 7, kmap=kmap)
    assert_close("a", a, ea, np.zeros_like(ta), ta)
    assert_close("b", b, eb, np.zeros_like(tb), tb)
    assert_close("c", c, ec, np.zeros_like(tc), tc)


def test_hessian_near_cores(hip, oracle):
    """Pixels at 1e-3..3 core radii of small-core dPIE and PIEMD halos, e = 0 (epot = 1e-5) halos."""
    rng = np.random.default_rng(3)
    halos = [dict(type=81, x=0.0, y=0.0, sigma=300.0, e=0.0, theta=0.4, rcore=0.05, rcut=20.0),
             dict(type=81, x=0.37, y=-0.21, sigma=200.0, e=0.45, theta=1.9, rcore=0.02, rcut=8.0),
             dict(type=8, x=-0.3, y=0.25, sigma=150.0, e=0.3, theta=2.5, rcore=0.1)]
    lens = M.LensModel.from_halos(halos)
    grid = M.Grid(M.grid_param(-0.6, 0.6, -0.55, 0.55), 211, 197)
    a, b, c, d = hip.hessian_grid(lens, grid)
    (ea, eb, ec), (ta, tb, tc) = hessian_reference(oracle, lens, grid)
    assert_close("a", a, ea, np.zeros_like(ta), ta)
    assert_close("b", b, eb, np.zeros_like(tb), tb)
    assert_close("c", c, ec, np.zeros_like(tc), tc)


def test_px_variants_identical(hip):
    """Pixels per thread changes the launch shape, never the per-pixel arithmetic."""
    lens = mixed_lens(n14=1)
    grid = M.Grid(M.grid_param(-40, 40, -35, 35), 300, 37)
    m = hip.Model(lens, hip.MAP_HESSIAN)
    s = hip.stream_create()
    res = []
    for px in (1, 2, 4):
        outs = [hip.DeviceArray(grid.npix) for _ in range(4)]
        m.map(hip.MAP_HESSIAN, 0, [o.ptr for o in outs], grid, px=px, stream=s)
        hip.sync(s)
        res.append([o.download() for o in outs])
    hip.stream_destroy(s)
    for k in range(4):
        np.testing.assert_array_equal(res[0][k], res[1][k])
        np.testing.assert_array_equal(res[0][k], res[2][k])
    host = hip.hessian_grid(lens, grid)
    for k in range(4):
        np.testing.assert_array_equal(host[k].ravel(), res[1][k])


def test_empty_lens_and_pitch(hip):
    lens = mixed_lens(n81=0, n8=0, n5=0)
    grid = M.Grid(M.grid_param(-1, 1, -1, 1), 70, 9)
    a, b, c, d = hip.hessian_grid(lens, grid)
    assert not a.any() and not b.any() and not c.any()
    # a pitched device map leaves the padding untouched
    lens = mixed_lens()
    m = hip.Model(lens, hip.MAP_HESSIAN)
    pitch = 96
    outs = [hip.DeviceArray(grid.ny * pitch) for _ in range(4)]
    for o in outs:
        o.upload(np.full(grid.ny * pitch, 7.0))
    m.map(hip.MAP_HESSIAN, 0, [o.ptr for o in outs], grid, pitch=pitch)
    hip.sync()
    ref = hip.hessian_grid(lens, grid)
    for k in range(4):
        v = outs[k].download().reshape(grid.ny, pitch)
        np.testing.assert_array_equal(v[:, :grid.nx], ref[k])
        assert (v[:, grid.nx:] == 7.0).all()


# ------------------------------------------------------------------------------------------------
# maps
# ------------------------------------------------------------------------------------------------

def _epilogue_reference(oracle, kind, mode, lens, grid, cosmo, zl, zs, kmap=False):
    ea, eb, ec, ed = oracle.hessian_planes(lens, grid, kmap=kmap, extended=True)
    dl0s = oracle.cosmo_object_object(zl, zs, cosmo)
    dos = oracle.cosmo_observer_object(zs, cosmo)
    dol = oracle.cosmo_observer_object(zl, cosmo)
    (out,) = oracle.map_epilogue(kind, mode, ea, eb, ec, dl0s, dos, dol, cosmo.h, grid.dx, grid.dy)
    return out.reshape(grid.ny, grid.nx)


@pytest.mark.parametrize("kind,mode", [(1, m) for m in range(1, 7)] + [(2, m) for m in range(1, 5)] +
                         [(3, m) for m in range(1, 5)])
def test_maps_reference_modes(hip, oracle, kind, mode):
    lens = M.group_lens(seed=4)
    grid = M.group_grid(96)
    cosmo = M.flat_lcdm()
    zl, zs = 0.4, 2.0
    if kind == 3:
        lens.dlsds[:] = oracle.cosmo_dlsds(zl, 1.5, cosmo)
        lens._soa = None
    got = hip.map_grid(kind, mode, lens, grid, cosmo, zl if kind != 3 else 0.0, zs)
    want = _epilogue_reference(oracle, kind, mode, lens, grid, cosmo, zl, zs, kmap=(kind == 3))
    # the maps are smooth functions of (a, b, c): compare relative to the map's own scale, with a
    # looser bound where det A ~ 0 (amplification diverges on the critical curves)
    (ea, eb, ec), (ta, tb, tc) = hessian_reference(oracle, lens, grid, kmap=(kind == 3))
    r = oracle.cosmo_dlsds(zl, zs, cosmo)
    A, B, Cc = 1 - ea * r, -eb * r, 1 - ec * r
    det = A * Cc - B * B
    tolH = np.maximum(np.maximum(ta, tb), tc) * max(r, 1.0)
    if kind == 1 and mode in (1, 2, 4):
        tol = 8 * tolH / np.maximum(det * det, 1e-300) * (np.abs(A) + np.abs(B) + np.abs(Cc) + 1)
    elif kind == 1 and mode == 3:
        tol = 8 * tolH / np.maximum(np.abs(det), 1e-300) * (np.abs(A) + np.abs(B) + np.abs(Cc) + 1)
    elif kind == 2 and mode == 2:
        gam = np.sqrt(((A - Cc) / 2) ** 2 + B ** 2)
        lam, mu = (A + Cc) / 2 + gam, (A + Cc) / 2 - gam
        tol = 16 * tolH * (np.abs(lam) + np.abs(mu) + 1) ** 2 / np.maximum(lam * lam + mu * mu, 1e-300) ** 1.5
    else:
        tol = 8 * tolH * max(1.0, np.abs(want).max())
    ok = np.abs(got - want) <= tol + 1e-14 * np.abs(want)
    assert ok.mean() == 1.0, (kind, mode, np.abs(got - want)[~ok][:5], tol[~ok][:5])


def test_map_fields_and_dpl(hip, oracle):
    lens = M.group_lens(seed=4)
    grid = M.group_grid(96)
    cosmo = M.flat_lcdm()
    zl, zs = 0.4, 2.0
    kap, g1, g2, det, mu = hip.map_fields(lens, grid, cosmo, zl, zs)
    (ea, eb, ec), (ta, tb, tc) = hessian_reference(oracle, lens, grid)
    r = oracle.cosmo_dlsds(zl, zs, cosmo)
    ek, eg1, eg2 = (ea + ec) * r / 2, (ea - ec) * r / 2, eb * r
    edet = (1 - ek) ** 2 - eg1 ** 2 - eg2 ** 2
    tk = (ta + tc) * r
    assert_close("kappa", kap, ek, np.zeros_like(tk), tk + 1e-15)
    assert_close("gamma1", g1, eg1, np.zeros_like(tk), tk + 1e-15)
    assert_close("gamma2", g2, eg2, np.zeros_like(tk), tb * r + 1e-15)
    tdet = 4 * (np.abs(1 - ek) + np.abs(eg1) + np.abs(eg2)) * (tk + tb * r) + 1e-15
    assert_close("detA", det, edet, np.zeros_like(tdet), tdet)
    np.testing.assert_array_equal(mu, 1.0 / det)
    dx, dy = hip.map_grid(hip.MAP_DPL, 1, lens, grid, cosmo, zl, zs)
    gx, gy = hip.gradient_grid(lens, grid)
    np.testing.assert_allclose(dx, gx * r, rtol=1e-15, atol=0)
    np.testing.assert_allclose(dy, gy * r, rtol=1e-15, atol=0)


@pytest.mark.parametrize("htype", [81, 8])
def test_potential_next_to_the_centre(hip, oracle, htype):
    """Pixels 1e-3 to 3e-2 core radii from a halo centre (never on it): the potential's logarithms
    are formed as log1p of small quantities, so the 1e-12 criterion holds there too
    (written as the logarithm of a quotient of nearly equal numbers it lost up to 1e-8 of the
    potential at such pixels)."""
    halos = [dict(type=htype, x=0.37, y=-0.21, sigma=420.0, e=0.35, theta=0.6, rcore=0.5, rcut=60.0),
             dict(type=81, x=8.0, y=5.0, sigma=250.0, e=0.2, theta=1.9, rcore=0.2, rcut=30.0)]
    lens = M.LensModel.from_halos(halos)
    n, step = 33, 1e-3
    x0, y0 = 0.37 - (n // 2 + 0.5) * step, -0.21 - (n // 2 + 0.5) * step
    grid = M.Grid(M.grid_param(x0, x0 + (n - 1) * step, y0, y0 + (n - 1) * step), n, n)
    print_case(f"potential next to a type-{htype} centre", lens, grid)
    phi = hip.potential_grid(lens, grid)
    ext = oracle.potential_grid(lens, grid, extended=True)
    S = np.zeros_like(ext)
    for i in range(lens.n):
        S += np.abs(oracle.potential_grid(lens.subset([i]), grid, extended=True))
    assert np.isfinite(phi).all()
    r = assert_close("phi", phi, ext, S, np.zeros_like(S))
    print(f"[potential next to a type-{htype} centre] max error / tolerance {r:.3e}", flush=True)


def test_potential_nan_exactly_on_a_centre(hip, oracle):
    """A pixel exactly on a PIEMD or dPIE centre: the reference's potential is 0/0 there (NaN, as the
    oracle's restatement gives); the kernel's form has the limit 0 there and returns NaN instead,
    at that pixel only. The grid passes through both centres; an SIS centre stays finite (b0 R = 0)."""
    halos = [dict(type=81, x=1.0, y=-2.0, sigma=500.0, e=0.3, theta=0.4, rcore=0.3, rcut=40.0),
             dict(type=8, x=-3.0, y=2.0, sigma=300.0, e=0.2, theta=1.1, rcore=0.5, rcut=30.0),
             dict(type=5, x=2.0, y=2.0, sigma=200.0, e=0.0, theta=0.0, rcore=0.0, rcut=0.0)]
    lens = M.LensModel.from_halos(halos)
    grid = M.Grid(M.grid_param(-4.0, 4.0, -4.0, 4.0), 9, 9)  # pixels on every integer point
    print_case("potential on halo centres", lens, grid)
    phi = hip.potential_grid(lens, grid)
    ext = oracle.potential_grid(lens, grid, extended=True)
    nan = ~np.isfinite(phi)
    assert nan.sum() == 2 and nan[2, 5] and nan[6, 1]  # rows y = -2 and y = 2, columns x = 1 and x = -3
    np.testing.assert_array_equal(~np.isfinite(ext), nan)
    S = np.zeros_like(ext)
    for i in range(lens.n):
        S += np.abs(oracle.potential_grid(lens.subset([i]), grid, extended=True))
    fin = ~nan
    assert_close("phi", phi[fin], ext[fin], S[fin], np.zeros(fin.sum()))


@pytest.mark.parametrize("epot", [0.3, 0.6])
def test_potential_lone_dpie_far_and_near(hip, oracle, epot):
    """One dPIE alone (so the criterion is 1e-12 of its own |phi|): pixels 30 to 300 cut radii out,
    where the core and cut radii's deflections nearly cancel, and pixels 0.02 to 0.5 core radii from
    the centre, where every term vanishes; both merged parts are formed without that cancellation."""
    lens = M.LensModel(type=np.array([81], np.int32), x=np.array([0.4]), y=np.array([-0.3]),
                       b0=np.array([12.0]), angle=np.array([0.7]), epot=np.array([epot]),
                       rcore=np.array([0.2]), rcut=np.array([1.5]), z=np.array([0.4]), dlsds=np.array([1.0]))
    for lo, hi in ((45.0, 450.0), (-450.0, -45.0)):
        grid = M.Grid(M.grid_param(lo, hi, lo, hi), 48, 40)
        print_case(f"lone dPIE e={epot} far {lo}..{hi}", lens, grid)
        phi = hip.potential_grid(lens, grid)
        ext = oracle.potential_grid(lens, grid, extended=True)
        r = assert_close("phi far", phi, ext, np.abs(ext), np.zeros_like(ext))
        print(f"[lone dPIE e={epot} far] max error / tolerance {r:.3e}", flush=True)
    x0, y0, step, n = 0.4 + 0.004, -0.3 + 0.004, 0.1 / 40, 40
    grid = M.Grid(M.grid_param(x0, x0 + (n - 1) * step, y0, y0 + (n - 1) * step), n, n)
    print_case(f"lone dPIE e={epot} near the centre", lens, grid)
    phi = hip.potential_grid(lens, grid)
    ext = oracle.potential_grid(lens, grid, extended=True)
    r = assert_close("phi near", phi, ext, np.abs(ext), np.zeros_like(ext))
    print(f"[lone dPIE e={epot} near] max error / tolerance {r:.3e}", flush=True)


@pytest.mark.parametrize("which", ["mixed", "cluster"])
def test_potential(hip, oracle, which):
    """The potential (the deflection form of the reference's pi05) against the extended oracle of the
    reference's complex form, held to the FP64 criterion: 1e-12 of the sum over halos of |phi_h| at
    each pixel (the GPU's own error is below 1e-14 of it; no cancellation term is needed)."""
    if which == "mixed":
        lens = mixed_lens(n81=3, n8=2, n5=2)
        grid = M.Grid(M.grid_param(-40, 40, -35, 35), 64, 48)
    else:
        lens = M.cluster_lens(seed=3, ngal=60)
        grid = M.Grid(M.grid_param(-150, 150, -150, 150), 96, 80)
    print_case(f"potential {which}", lens, grid)
    phi = hip.potential_grid(lens, grid)
    ext = oracle.potential_grid(lens, grid, extended=True)
    S = np.zeros_like(ext)
    for i in range(lens.n):
        S += np.abs(oracle.potential_grid(lens.subset([i]), grid, extended=True))
    assert np.isfinite(phi).all()
    r = assert_close("phi", phi, ext, S, np.zeros_like(S))
    print(f"[potential {which}] max error / tolerance {r:.3e}", flush=True)
    # 1, 2 and 4 pixels per thread give the same map bit for bit
    m = hip.Model(lens, hip.MAP_POT)
    out = hip.DeviceArray(grid.npix)
    maps = []
    for px in (1, 2, 4):
        m.map(hip.MAP_POT, 3, [out.ptr], grid, px=px)
        hip.sync()
        maps.append(out.download().reshape(grid.ny, grid.nx))
    m.close()
    out.close()
    np.testing.assert_array_equal(maps[0], phi)
    np.testing.assert_array_equal(maps[1], phi)
    np.testing.assert_array_equal(maps[2], phi)


# ------------------------------------------------------------------------------------------------
# full-size (4096^2, 503 halos) properties
# ------------------------------------------------------------------------------------------------

def test_cluster_fullsize_sampled(hip, oracle):
    """The cluster Hessian workload at full size (4096^2, 503 dPIE): the GPU grid at 2000 random
    pixels and along whole rows through the 3 cluster-scale cores and 50 galaxy cores equals the
    extended oracle, and rows are independent of the grid they are computed in (a sub-grid with
    istart/jstart offsets reproduces the same pixels bit for bit)."""
    lens = M.cluster_lens()
    grid = M.cluster_grid(4096)
    print_case("cluster lens full size", lens, grid)
    a, b, c, d = hip.hessian_grid(lens, grid)
    assert np.isfinite(a).all() and np.isfinite(c).all()
    rng = np.random.default_rng(1)
    rows = rng.integers(0, grid.ny, 2000)
    cols = rng.integers(0, grid.nx, 2000)
    worst = 0.0
    for rr, cc in zip(rows, cols):
        px = grid.frame.xmin + cc * grid.dx
        py = grid.frame.ymin + rr * grid.dy
        e = oracle.point_hessian(lens, px, py, extended=True)
        s = oracle.point_hessian_abs(lens, px, py)
        _, ch = oracle.point_cond(lens, px, py)
        tol = REL * max(s[:3]) + EXT_SHARE * ch
        err = max(abs(a[rr, cc] - e[0]), abs(b[rr, cc] - e[1]), abs(c[rr, cc] - e[2]))
        worst = max(worst, err / tol)
    assert worst <= 1.0, worst
    # whole rows through the three cluster-scale cores and 50 galaxy cores, every pixel against the
    # extended oracle with the FP64 criterion
    big3 = np.argsort(lens.b0)[-3:]
    gal = np.setdiff1d(np.arange(lens.n), big3)
    pick = np.concatenate([big3, rng.choice(gal, 50, replace=False)])
    rows_sel = sorted({int(round((lens.y[h] - grid.frame.ymin) / grid.dy)) for h in pick
                       if 0 <= round((lens.y[h] - grid.frame.ymin) / grid.dy) < grid.ny})
    line = M.Grid(grid.frame, grid.nx, 1, step=(grid.dx, grid.dy))
    worst = 0.0
    for j in rows_sel:
        (ea, eb, ec), (ta, tb, tc) = hessian_reference(oracle, lens, line, 0, j)
        for name, got, ex, tol in (("a", a, ea, ta), ("b", b, eb, tb), ("c", c, ec, tc)):
            worst = max(worst, assert_close(f"{name} row {j}", got[j:j + 1], ex, np.zeros_like(tol), tol))
    assert len(rows_sel) >= 45 and worst <= 1.0
    # sub-grid through istart/jstart
    sub = M.Grid(grid.frame, 300, 200, step=(grid.dx, grid.dy))
    sa, sb, sc, sd = hip.hessian_grid(lens, sub, 1500, 2100)
    np.testing.assert_array_equal(sa, a[2100:2300, 1500:1800])
    np.testing.assert_array_equal(sc, c[2100:2300, 1500:1800])


# ------------------------------------------------------------------------------------------------
# a real model: the A2744 parameter file read by the library's reader (258 dPIE halos)
# ------------------------------------------------------------------------------------------------

def test_real_model_a2744(hip, oracle):
    """Hessian and gradient of the 258-halo A2744 model (6 cluster-scale dPIE + 252 potfile galaxies)
    on a 200x160 grid over the cluster core, against the oracle with the FP64 criterion, and the
    FP32C det A within 1e-6 of FP64."""
    import os
    from conftest import GOLDEN
    d = hip.load_par(os.path.join(GOLDEN, "par", "A2744_SLspec.par"))
    lens = d["lens"]
    grid = M.Grid(M.grid_param(-60.0, 60.0, -50.0, 46.0), 200, 160)
    print_case("A2744 real model", lens, grid)
    a, b, c, _ = hip.hessian_grid(lens, grid)
    ea, eb, ec, _ = oracle.hessian_planes(lens, grid, extended=True)
    _, _, _, _, aa, ab, ac = oracle.hessian_planes(lens, grid, with_abs=True)
    _, ch = oracle.cond_planes(lens, grid)
    for got, ex, sc in ((a, ea, aa), (b, eb, ab), (c, ec, ac)):
        assert (np.abs(got - ex) <= REL * sc + EXT_SHARE * ch).all()
    gx, gy = hip.gradient_grid(lens, grid)
    ex, ey = oracle.gradient_grid(lens, grid, extended=True)
    _, _, ax, ay = oracle.gradient_grid(lens, grid, with_abs=True)
    cg, _ = oracle.cond_planes(lens, grid)
    assert (np.abs(gx - ex) <= REL * ax + EXT_SHARE * cg).all() and (np.abs(gy - ey) <= REL * ay + EXT_SHARE * cg).all()
    r = hip.cosmo_dlsds(0.308, 2.0, d["cosmo"])
    f64 = hip.map_fields(lens, grid, d["cosmo"], 0.308, 2.0)
    f32 = hip.map_fields(lens, grid, d["cosmo"], 0.308, 2.0, precision=hip.PREC_FP32C)
    assert np.abs(f32[3] - f64[3]).max() <= 1e-6
    assert r > 0
