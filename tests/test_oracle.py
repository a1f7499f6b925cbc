"""The FP64 oracle against what pins it: the reference's known-answer test and golden vectors from
the reference's own compiled CPU functions, finite differences, closed forms and invariants."""
import json
import math
import os

import numpy as np
import pytest

import ltmodel as M
from conftest import GOLDEN


def lens_from_npz(z):
    n = len(z["type"])
    return M.LensModel(type=z["type"], x=z["x"], y=z["y"], b0=z["b0"], angle=z["angle"], epot=z["epot"],
                       rcore=z["rcore"], rcut=z["rcut"], z=np.full(n, 0.4), dlsds=np.ones(n))


def mixed_lens(seed=7, n81=4, n8=3, n5=4):
    rng = np.random.default_rng(seed)
    halos = []
    for t, k in ((81, n81), (8, n8), (5, n5)):
        for _ in range(k):
            halos.append(dict(type=t, x=float(rng.uniform(-30, 30)), y=float(rng.uniform(-30, 30)),
                              sigma=float(rng.uniform(150, 700)), e=float(rng.uniform(0.05, 0.6)),
                              theta=float(rng.uniform(0, math.pi)), rcore=float(rng.uniform(0.2, 5)),
                              rcut=float(rng.uniform(20, 300))))
    return M.LensModel.from_halos(halos)


# --------------------------------------------------------------------------------------------------
# reference known answer and golden vectors
# --------------------------------------------------------------------------------------------------

def test_kat1_gradient(oracle):
    """KAT-1: Benchmarks/GradientBenchmark/setup.cpp:19-44 (95 dPIE halos, one point)."""
    d = json.load(open(os.path.join(GOLDEN, "kat1_gradient.json")))
    n = len(d["type"])
    L = M.LensModel(type=d["type"], x=d["pos_x"], y=d["pos_y"], b0=d["b0"], angle=d["theta"],
                    epot=d["epot"], rcore=d["rcore"], rcut=d["rcut"], z=np.zeros(n), dlsds=np.ones(n))
    gx, gy = oracle.point_gradient(L, d["x"], d["y"])
    assert abs(gx - d["grad_x"]) <= 1e-13 * abs(d["grad_x"])
    assert abs(gy - d["grad_y"]) <= 1e-13 * abs(d["grad_y"])


def test_hessian_grid_golden(oracle):
    """dPIE Hessian grid of the compiled reference gradient2_grid_CPU (with istart/jstart).

    The reference and the double oracle both sit within the FP64 rounding scale of the closed form
    (oracle cond planes) of the extended-precision oracle, and within 1e-13 of the sum of |terms|
    except where that scale is larger (next to the removable 0/0 point of the PIEMD terms)."""
    z = np.load(os.path.join(GOLDEN, "hess81_grid.npz"))
    L = lens_from_npz(z)
    nx, ny, i0, j0 = (int(v) for v in z["shape"])
    grid = M.Grid(M.grid_param(*z["frame"]), nx, ny, step=tuple(z["dxdy"]))
    a, b, c, d, aa, ab, ac = oracle.hessian_planes(L, grid, i0, j0, with_abs=True)
    xa, xb, xc, xd = oracle.hessian_planes(L, grid, i0, j0, extended=True)
    _, ch = oracle.cond_planes(L, grid, i0, j0)
    for got, want, ext, scale in ((a, z["a"], xa, aa), (b, z["b"], xb, ab), (c, z["c"], xc, ac),
                                  (d, z["d"], xd, ab)):
        tol = 1e-13 * scale + ch
        assert (np.abs(want - ext) <= tol).all()      # the reference itself
        assert (np.abs(got - ext) <= tol).all()       # the double restatement
        assert (np.abs(got - want) <= 2 * tol).all()
        # most pixels agree with the reference to 1e-13 of the term sum outright
        assert np.mean(np.abs(got - want) <= 1e-13 * scale) > 0.98


def test_hessian_points_golden(oracle):
    z = np.load(os.path.join(GOLDEN, "hess81_points.npz"))
    L = lens_from_npz(z)
    for (px, py), h in zip(z["points"], z["hessian"]):
        got = np.array(oracle.point_hessian(L, px, py))
        ext = np.array(oracle.point_hessian(L, px, py, extended=True))
        _, ch = oracle.point_cond(L, px, py)
        tol = 1e-13 * np.abs(h).max() + ch
        assert np.abs(h - ext).max() <= tol
        assert np.abs(got - ext).max() <= tol


def test_delense_golden(oracle):
    z = np.load(os.path.join(GOLDEN, "delense.npz"))
    n = int(z["n"][0])
    frame = M.grid_param(*z["frame"])
    src = np.ascontiguousarray(z["sources"])
    st, pos, loc, numimg = oracle.delense_barycenter(len(src), n, frame, src, z["gx"], z["gy"])
    assert st == 0
    np.testing.assert_array_equal(loc, z["locimagesfound"])
    assert numimg == int(z["numimg"][0])
    for s in range(len(src)):
        k = loc[s]
        np.testing.assert_allclose(pos[s, :k], z["image_pos"][s, :k], rtol=0, atol=1e-12)


def test_chi_golden(oracle):
    z = np.load(os.path.join(GOLDEN, "chi.npz"))
    images = np.ascontiguousarray(z["images"])
    chi, found = oracle.chi_computation(z["locimagesfound"], z["image_pos"], z["nimages"], images)
    assert found == int(z["images_found"][0])
    assert abs(chi - z["chi"][0]) <= 1e-12 * abs(z["chi"][0])


def test_oracle_matches_reference_build_random(oracle):
    """Fresh random dPIE lenses through the reference build, when it is present (this container)."""
    R = oracle.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built (no reference sources on this machine)")
    import ctypes as C
    rng = np.random.default_rng(99)
    for trial in range(3):
        halos = [dict(type=81, x=float(rng.uniform(-10, 10)), y=float(rng.uniform(-10, 10)),
                      sigma=float(rng.uniform(100, 1000)), e=float(rng.uniform(0, 0.8)),
                      theta=float(rng.uniform(0, 3.2)), rcore=float(10 ** rng.uniform(-1.5, 1)),
                      rcut=float(10 ** rng.uniform(0.5, 3))) for _ in range(6)]
        L = M.LensModel.from_halos(halos)
        grid = M.Grid(M.grid_param(-12, 12, -12, 12), 17, 23)
        out = (M.LtMatrix * grid.npix)()
        R.ref_gradient2_grid(out, C.byref(grid.frame), C.byref(L.soa()), L.n, grid.dx, grid.dy,
                             grid.nx, grid.ny, 0, 0)
        ref = np.ctypeslib.as_array(out).view(np.float64).reshape(grid.ny, grid.nx, 4)
        a, b, c, d, aa, ab, ac = oracle.hessian_planes(L, grid, with_abs=True)
        ext = oracle.hessian_planes(L, grid, extended=True)
        _, ch = oracle.cond_planes(L, grid)
        for k, (got, sc) in enumerate(((a, aa), (b, ab), (c, ac), (d, ab))):
            tol = 1e-13 * sc + ch
            assert (np.abs(ref[..., k] - ext[k]) <= tol).all()
            assert (np.abs(got - ext[k]) <= tol).all()


# --------------------------------------------------------------------------------------------------
# mathematics: finite differences, limits, invariants
# --------------------------------------------------------------------------------------------------

def _fd_points(L, rng, k=12):
    pts = []
    for _ in range(k):
        i = rng.integers(L.n)
        r = rng.uniform(0.5, 4.0) * max(L.rcore[i], 0.5)
        a = rng.uniform(0, 2 * math.pi)
        pts.append((L.x[i] + r * math.cos(a), L.y[i] + r * math.sin(a)))
    return pts


@pytest.mark.parametrize("t", [5, 8, 81])
def test_gradient_is_grad_of_potential(oracle, t):
    L = mixed_lens()
    L = L.subset(np.nonzero(L.type == t)[0])
    rng = np.random.default_rng(t)
    h = 1e-4
    for px, py in _fd_points(L, rng):
        gx, gy = oracle.point_gradient(L, px, py)
        fx = (oracle.point_potential(L, px + h, py) - oracle.point_potential(L, px - h, py)) / (2 * h)
        fy = (oracle.point_potential(L, px, py + h) - oracle.point_potential(L, px, py - h)) / (2 * h)
        scale = abs(gx) + abs(gy)
        assert abs(fx - gx) < 1e-7 * scale and abs(fy - gy) < 1e-7 * scale


@pytest.mark.parametrize("t", [5, 8, 81])
def test_hessian_is_grad_of_gradient(oracle, t):
    L = mixed_lens()
    L = L.subset(np.nonzero(L.type == t)[0])
    rng = np.random.default_rng(10 + t)
    h = 1e-5
    for px, py in _fd_points(L, rng):
        a, b, c, d = oracle.point_hessian(L, px, py)
        assert abs(b - d) <= 1e-14 * (abs(a) + abs(b) + abs(c))
        gxp, gyp = oracle.point_gradient(L, px + h, py)
        gxm, gym = oracle.point_gradient(L, px - h, py)
        hx = ((gxp - gxm) / (2 * h), (gyp - gym) / (2 * h))
        gxp, gyp = oracle.point_gradient(L, px, py + h)
        gxm, gym = oracle.point_gradient(L, px, py - h)
        hy = ((gxp - gxm) / (2 * h), (gyp - gym) / (2 * h))
        scale = abs(a) + abs(b) + abs(c)
        assert abs(hx[0] - a) < 1e-6 * scale
        assert abs(hy[1] - c) < 1e-6 * scale
        assert abs(hx[1] - b) < 1e-6 * scale and abs(hy[0] - b) < 1e-6 * scale


def test_piemd81_circular_limit(oracle):
    """eps -> 0: dPIE deflection -> b0 rcut/(rcut-rc) (sqrt(u+rc^2) - rc - sqrt(u+rcut^2) + rcut)/u
    (x, y)  (Eliasdottir et al. 2007 Eq. A23)."""
    rc, rcut, b0 = 2.0, 80.0, 10.0
    L = M.LensModel.from_halos([dict(type=81, x=0.0, y=0.0, b0=b0, epot=1e-9, theta=0.3, rcore=rc,
                                     rcut=rcut)])
    for px, py in ((3.0, 1.0), (-20.0, 7.0), (0.5, -0.2), (150.0, 90.0)):
        u = px * px + py * py
        t = (math.sqrt(u + rc * rc) - rc - math.sqrt(u + rcut * rcut) + rcut) * b0 * rcut / (rcut - rc) / u
        gx, gy = oracle.point_gradient(L, px, py)
        assert abs(gx - t * px) < 1e-5 * abs(t) * math.hypot(px, py)
        assert abs(gy - t * py) < 1e-5 * abs(t) * math.hypot(px, py)


def test_piemd8_is_large_rcut_limit_of_81(oracle):
    base = dict(x=1.0, y=-2.0, b0=5.0, epot=0.2, theta=0.7, rcore=1.5)
    L8 = M.LensModel.from_halos([dict(type=8, **base)])
    L81 = M.LensModel.from_halos([dict(type=81, rcut=1e9, **base)])
    for px, py in ((3.0, 1.0), (-8.0, 4.0), (0.3, 0.1)):
        g8 = np.array(oracle.point_gradient(L8, px, py))
        g81 = np.array(oracle.point_gradient(L81, px, py))
        assert np.abs(g8 - g81).max() < 1e-6 * np.abs(g8).max()
        h8 = np.array(oracle.point_hessian(L8, px, py))
        h81 = np.array(oracle.point_hessian(L81, px, py))
        assert np.abs(h8 - h81).max() < 1e-6 * np.abs(h8).max()


def test_shear14_invariants(oracle):
    for eps, th in ((0.05, 0.3), (0.12, 2.0)):
        L = M.LensModel.from_halos([dict(type=14, x=0.0, y=0.0, b0=0.0, epot=eps, theta=th)])
        a, b, c, d = oracle.point_hessian(L, 3.0, -4.0)
        assert abs(a + c) < 1e-15                       # traceless
        assert abs((a * c - b * d) + 9 * eps * eps) < 1e-14   # eigenvalues +-3 eps
        assert b == d


def test_kmap_weights_runs(oracle):
    L = mixed_lens(n81=3, n8=0, n5=3)
    L.dlsds[:] = np.where(L.type == 5, 0.6, 0.85)
    L._soa = None
    Ls = L.subset(np.nonzero(L.type == 5)[0])
    Lp = L.subset(np.nonzero(L.type == 81)[0])
    for px, py in ((1.0, 2.0), (-7.0, 3.5)):
        k = np.array(oracle.point_hessian(L, px, py, kmap=True))
        want = 0.6 * np.array(oracle.point_hessian(Ls, px, py)) + 0.85 * np.array(oracle.point_hessian(Lp, px, py))
        assert np.abs(k - want).max() < 1e-13 * np.abs(want).max()


def test_epilogue_closed_forms(oracle):
    rng = np.random.default_rng(5)
    a, b, c = rng.normal(0.3, 0.4, 50), rng.normal(0, 0.2, 50), rng.normal(0.3, 0.4, 50)
    dl0s, dos, dol, h = 0.7, 0.9, 0.5, 1.4
    r = dl0s / dos
    kappa = (a + c) * r / 2
    g1, g2 = (a - c) * r / 2, b * r
    gam = np.hypot(g1, g2)
    det = (1 - kappa) ** 2 - gam ** 2
    # Jacobian eigenvalues are 1 - kappa -+ gamma
    (mu1,) = oracle.map_epilogue(1, 1, a, b, c, dl0s, dos, dol, h, 0.1, 0.1)
    np.testing.assert_allclose(mu1, 1 / det, rtol=1e-9)
    (s1,) = oracle.map_epilogue(2, 1, a, b, c, dl0s, dos, dol, h, 0.1, 0.1)
    np.testing.assert_allclose(s1, gam, rtol=1e-12, atol=1e-15)
    (k5,) = oracle.map_epilogue(1, 5, a, b, c, dl0s, dos, dol, h, 0.1, 0.1)
    np.testing.assert_allclose(k5, (a + c) / 2, rtol=1e-15)
    outs = oracle.map_epilogue(6, 0, a, b, c, dl0s, dos, dol, h, 0.1, 0.1, nout=5)
    for got, want in zip(outs, (kappa, g1, g2, det, 1 / det)):
        np.testing.assert_allclose(got, want, rtol=1e-12)
    (m2,) = oracle.map_epilogue(3, 2, a, b, c, dl0s, dos, dol, h, 0.1, 0.1)
    np.testing.assert_allclose(m2, (a + c) / 2 * (0.057929405 * h / dol / dl0s * dos), rtol=1e-14)


def test_cosmology_quadrature(oracle):
    from scipy.integrate import quad
    c = M.flat_lcdm(70.0, 0.3)
    E = lambda z: math.sqrt(0.3 * (1 + z) ** 3 + 0.7)
    for zl, zs in ((0.4, 2.0), (0.3, 5.7), (0.1, 1.0)):
        dls = quad(lambda z: 1 / E(z), zl, zs, epsabs=1e-15, epsrel=1e-13)[0] / (1 + zs)
        ds = quad(lambda z: 1 / E(z), 0, zs, epsabs=1e-15, epsrel=1e-13)[0] / (1 + zs)
        assert abs(oracle.cosmo_dlsds(zl, zs, c) - dls / ds) < 1e-13
    # Omega_X = 0: Mattig closed form against the integral of the matter-only (+ curvature) model
    c0 = M.flat_lcdm(70.0, 0.3)
    c0.omegaX = 0.0
    c0.curvature = 0.7   # open, Omega_k = 1 - Omega_m; not used by the Mattig branch
    zs = 2.0
    Ok = 0.7
    chi = quad(lambda z: 1 / math.sqrt(0.3 * (1 + z) ** 3 + Ok * (1 + z) ** 2), 0, zs, epsabs=1e-15, epsrel=1e-13)[0]
    da = math.sinh(math.sqrt(Ok) * chi) / math.sqrt(Ok) / (1 + zs)
    assert abs(oracle.cosmo_observer_object(zs, c0) - da) < 1e-12


def test_epilogue_mode_identities(oracle):
    """Ampli modes 2, 3, 4, 6 and shear modes 1-4 against quantities computed independently from the
    matrices: eigenvalues of the Jacobian J = 1 - r H and of H (numpy's symmetric eigensolver), and the
    spin-2 behaviour of (gamma1, gamma2) when the frame rotates (H' = R^T H R by matrix products)."""
    rng = np.random.default_rng(11)
    n = 200
    a, b, c = rng.normal(0.3, 0.5, n), rng.normal(0, 0.3, n), rng.normal(0.3, 0.5, n)
    dl0s, dos, dol, h = 0.7, 0.9, 0.5, 1.4
    r = dl0s / dos
    H = np.stack([np.stack([a, b], -1), np.stack([b, c], -1)], -2)
    J = np.eye(2) - r * H
    lj = np.linalg.eigvalsh(J)            # ascending: mu <= lambda
    mu_j, lam_j = lj[:, 0], lj[:, 1]
    lh = np.linalg.eigvalsh(H)
    one_minus_kappa = 0.5 * (lam_j + mu_j)

    def ep(kind, mode, A=a, B=b, Cc=c):
        (v,) = oracle.map_epilogue(kind, mode, A, B, Cc, dl0s, dos, dol, h, 0.1, 0.1)
        return v

    np.testing.assert_allclose(ep(1, 2), 1 / np.abs(lam_j * mu_j), rtol=1e-9)
    np.testing.assert_allclose(ep(1, 3), -2.5 * np.log10(np.abs(lam_j * mu_j)), rtol=1e-9, atol=1e-12)
    # mode 4: (1 - kappa)(1 + g^2)/(1 - g^2), g = gamma/(1 - kappa) = (1 - kappa)(lam^2 + mu^2)/(2 lam mu)
    np.testing.assert_allclose(ep(1, 4), one_minus_kappa * (lam_j ** 2 + mu_j ** 2) / (2 * lam_j * mu_j), rtol=1e-9)
    # mode 6: the shear without D_ls/D_s = half the spread of the Hessian's eigenvalues
    np.testing.assert_allclose(ep(1, 6), 0.5 * (lh[:, 1] - lh[:, 0]), rtol=1e-12, atol=1e-14)
    # shear 1: (lambda - mu) / 2 of J; shear 2: |lam^2 - mu^2| / (lam^2 + mu^2)
    g = ep(2, 1)
    np.testing.assert_allclose(g, 0.5 * (lam_j - mu_j), rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(ep(2, 2), np.abs(lam_j ** 2 - mu_j ** 2) / (lam_j ** 2 + mu_j ** 2), rtol=1e-10,
                               atol=1e-14)
    # shear 3, 4 = (gamma1, gamma2): |(g1, g2)| = shear 1, and a rotation of the frame by phi turns
    # (g1, g2) by -2 phi
    g1, g2 = ep(2, 3), ep(2, 4)
    np.testing.assert_allclose(np.hypot(g1, g2), g, rtol=1e-12, atol=1e-14)
    for phi in (0.3, 1.1, -2.0):
        R = np.array([[np.cos(phi), -np.sin(phi)], [np.sin(phi), np.cos(phi)]])
        Hr = R.T @ H @ R
        g1r, g2r = ep(2, 3, Hr[:, 0, 0], Hr[:, 0, 1], Hr[:, 1, 1]), ep(2, 4, Hr[:, 0, 0], Hr[:, 0, 1], Hr[:, 1, 1])
        np.testing.assert_allclose(g1r, g1 * np.cos(2 * phi) + g2 * np.sin(2 * phi), rtol=1e-10, atol=1e-13)
        np.testing.assert_allclose(g2r, -g1 * np.sin(2 * phi) + g2 * np.cos(2 * phi), rtol=1e-10, atol=1e-13)
    # the sign convention of gamma1: a Hessian with a > c has gamma1 > 0
    assert ep(2, 3, np.array([1.0]), np.array([0.0]), np.array([0.2]))[0] > 0
