"""Delensing, chi^2 and point gradients on the GPU against the oracle and the reference's golden
vectors (src/delense_CPU.cpp, src/chi_computation.cpp, src/chi_CPU_barycenter.cpp).

Delensing and the chi^2 matching are compare-and-count work on given gradient planes: bit-exact.
The chain that starts from a lens model goes through the FP64 gradient (equal to the oracle within
1e-12 relative), so its chi^2 is compared with a relative tolerance and its found-image counts
must be equal."""
import os

import numpy as np
import pytest

import ltmodel as M
from conftest import GOLDEN
from parity import EXT_SHARE

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lt():
    import lthip
    lthip.check(lthip.lib().lt_init(0))
    return lthip


def _case(msg):
    print(f"[case] {msg}", flush=True)


def test_delense_golden(lt, oracle):
    z = np.load(os.path.join(GOLDEN, "delense.npz"))
    n = int(z["n"][0])
    frame = M.grid_param(*z["frame"])
    src = np.ascontiguousarray(z["sources"])
    _case(f"delense golden {len(src)} sources {n}x{n}")
    st, pos, loc, numimg = lt.delense_barycenter(len(src), n, frame, src, z["gx"], z["gy"])
    assert st == 0
    np.testing.assert_array_equal(loc, z["locimagesfound"])
    assert numimg == int(z["numimg"][0])
    ost, opos, oloc, onum = oracle.delense_barycenter(len(src), n, frame, src, z["gx"], z["gy"])
    np.testing.assert_array_equal(pos, opos)
    for s in range(len(src)):
        k = loc[s]
        np.testing.assert_allclose(pos[s, :k], z["image_pos"][s, :k], rtol=0, atol=1e-12)


def _folding_planes(n, periods, amp=30.0, dr=0.9):
    """Gradient planes whose source plane folds `periods` times per axis: s = A sin(k x), so one
    source has about (2 periods)^2 images."""
    xs = np.linspace(-100.0, 100.0, n)
    X, Y = np.meshgrid(xs, xs)
    k = 2 * np.pi * periods / 200.0
    return (X - amp * np.sin(k * X)) / dr, (Y - amp * np.sin(k * Y)) / dr


@pytest.mark.parametrize("periods", [3, 5])
def test_delense_overflow(lt, oracle, periods):
    """More than 20 images per source (36 and 100 here): status LT_ERR_OVERFLOW, the first 20 in
    scan order kept, every hit counted in numimg -- as the oracle does."""
    n = 64
    frame = M.grid_param(-100.0, 100.0, -100.0, 100.0)
    gx, gy = _folding_planes(n, periods)
    src = M.galaxies_array(3)
    src["cx"] = [0.3, 5.1, -12.0]
    src["cy"] = [0.7, -3.3, 8.0]
    src["dr"] = 0.9
    _case(f"delense overflow periods={periods}")
    st, pos, loc, numimg = lt.delense_barycenter(3, n, frame, src, gx.ravel(), gy.ravel())
    ost, opos, oloc, onum = oracle.delense_barycenter(3, n, frame, src, gx.ravel(), gy.ravel())
    assert ost == 5 and st == 5
    np.testing.assert_array_equal(loc, oloc)
    assert numimg == onum > 60
    np.testing.assert_array_equal(pos, opos)


def test_delense_no_hits(lt, oracle):
    """Sources outside the delensed grid: nothing found, status OK."""
    n = 32
    frame = M.grid_param(-10.0, 10.0, -10.0, 10.0)
    gx = np.zeros(n * n)
    gy = np.zeros(n * n)
    src = M.galaxies_array(2)
    src["cx"] = [50.0, -30.0]
    src["cy"] = [0.0, 40.0]
    src["dr"] = 1.0
    st, pos, loc, numimg = lt.delense_barycenter(2, n, frame, src, gx, gy)
    assert st == 0 and numimg == 0
    np.testing.assert_array_equal(loc, [0, 0])


def test_delense_degenerate_and_nonfinite(lt, oracle):
    """Cells the scan's cull must leave to the full reference test: a patch whose corners all map to
    one source-plane point (zero-area triangles, which the reference's border-inclusive test accepts
    for any source), sources on a mapped corner and on a cell edge, NaN and infinite gradients, a NaN
    source and an infinite D_ls/D_s. Found images, their order, counts and status equal the oracle's."""
    n = 48
    frame = M.grid_param(-10.0, 10.0, -10.0, 10.0)
    xs = np.linspace(-10.0, 10.0, n)
    X, Y = np.meshgrid(xs, xs)
    r = np.hypot(X, Y) + 0.3
    gx = 4.0 * X / r + 0.1 * Y
    gy = 4.0 * Y / r - 0.05 * X
    dr0, c = 0.8, (1.5, -2.0)
    gx[30:36, 5:12] = (X[30:36, 5:12] - c[0]) / dr0
    gy[30:36, 5:12] = (Y[30:36, 5:12] - c[1]) / dr0
    gx[10, 40] = np.nan
    gy[12, 20] = np.inf
    gx[40, 30] = -np.inf
    gx, gy = gx.ravel(), gy.ravel()
    src = M.galaxies_array(7)
    dr = 0.9
    j = 20 * n + 23  # a grid point away from the patches: the source sits exactly on its mapped corner
    src["cx"] = [c[0], 0.7, xs[23] - dr * gx[j], 0.5 * ((xs[23] - dr * gx[j]) + (xs[24] - dr * gx[j + 1])),
                 np.nan, 2.0, -3.0]
    src["cy"] = [c[1], -0.4, xs[20] - dr * gy[j], 0.5 * ((xs[20] - dr * gy[j]) + (xs[20] - dr * gy[j + 1])),
                 1.0, 0.5, 3.5]
    src["dr"] = [dr0, dr0, dr, dr, 0.9, np.inf, 0.7]
    _case("delense degenerate and non-finite cells")
    st, pos, loc, numimg = lt.delense_barycenter(7, n, frame, src, gx, gy)
    ost, opos, oloc, onum = oracle.delense_barycenter(7, n, frame, src, gx, gy)
    assert st == ost
    assert numimg == onum
    np.testing.assert_array_equal(loc, oloc)
    np.testing.assert_array_equal(pos, opos)
    assert loc[0] == 20 and loc[2] >= 1


def test_chi_computation_golden(lt, oracle):
    z = np.load(os.path.join(GOLDEN, "chi.npz"))
    images = np.ascontiguousarray(z["images"])
    _case("chi_computation golden")
    chi, found = lt.chi_computation(z["locimagesfound"], z["image_pos"], z["nimages"], images)
    ochi, ofound = oracle.chi_computation(z["locimagesfound"], z["image_pos"], z["nimages"], images)
    assert found == ofound == int(z["images_found"][0])
    assert chi == ochi
    assert abs(chi - z["chi"][0]) <= 1e-12 * abs(z["chi"][0])


def test_point_gradients(lt, oracle):
    lens = M.cluster_lens(seed=4, ngal=60)
    rng = np.random.default_rng(8)
    px = rng.uniform(-150, 150, 777)
    py = rng.uniform(-150, 150, 777)
    # and some points close to the cores
    px[:60] = lens.x[:60] + rng.normal(0, 0.3, 60)
    py[:60] = lens.y[:60] + rng.normal(0, 0.3, 60)
    _case("point gradients 777 points, 63 halos")
    gx, gy = lt.point_gradients(lens, px, py)
    for i in range(len(px)):
        e = oracle.point_gradient(lens, px[i], py[i], extended=True)
        cg, _ = oracle.point_cond(lens, px[i], py[i])
        sc = np.abs(np.array(oracle.point_gradient(lens, px[i], py[i]))).max() + 1.0
        assert abs(gx[i] - e[0]) <= 1e-12 * sc + EXT_SHARE * cg, i
        assert abs(gy[i] - e[1]) <= 1e-12 * sc + EXT_SHARE * cg, i


def _families(lens, oracle, n, frame, nsets, seed, per=3):
    """Observed image families from the oracle's own delensing of random sources (images = found
    ones + N(0, 0.3") noise, family order shuffled), as a strong-lensing catalogue looks."""
    rng = np.random.default_rng(seed)
    grid = M.Grid(frame, n, n)
    gx, gy = oracle.gradient_grid(lens, grid)
    src = M.galaxies_array(nsets)
    src["cx"] = rng.uniform(-6, 6, nsets)
    src["cy"] = rng.uniform(-6, 6, nsets)
    src["dr"] = rng.uniform(0.8, 0.95, nsets)
    _, pos, loc, _ = oracle.delense_barycenter(nsets, n, frame, src, gx.ravel(), gy.ravel())
    nimages = np.full(nsets, per, dtype=np.int32)
    images = M.galaxies_array(per * nsets)
    for s in range(nsets):
        for k in range(per):
            g = images[per * s + k]
            if k < loc[s]:
                g["cx"] = pos[s, k, 0] + rng.normal(0, 0.3)
                g["cy"] = pos[s, k, 1] + rng.normal(0, 0.3)
            else:
                g["cx"] = rng.uniform(-80, 80)
                g["cy"] = rng.uniform(-80, 80)
            g["dr"] = src["dr"][s]
            g["redshift"] = 1.5
            images[per * s + k] = g
        perm = rng.permutation(per)
        images[per * s:per * s + per] = images[per * s:per * s + per][perm]
    return nimages, images


def test_chi_barycentersource(lt, oracle):
    lens = M.group_lens(seed=5)
    n = 96
    frame = M.grid_param(-100.0, 100.0, -100.0, 100.0)
    nimages, images = _families(lens, oracle, n, frame, 12, seed=21)
    _case(f"chi_barycentersource group lens n={n} 12 families")
    st, chi = lt.chi_barycentersource(lens, n, frame, nimages, images)
    ost, ochi = oracle.chi_barycentersource(lens, n, frame, nimages, images)
    assert st == ost == 0
    assert chi > 0
    assert abs(chi - ochi) <= 1e-9 * abs(ochi)


def test_chi_batch_matches_single(lt, oracle):
    """A batch of perturbed lens samples in one call: each sample's chi^2 and delensed images equal
    the single-model path (bitwise) and the oracle (within the gradient tolerance)."""
    base = M.group_lens(seed=5)
    n = 80
    frame = M.grid_param(-100.0, 100.0, -100.0, 100.0)
    nimages, images = _families(base, oracle, n, frame, 9, seed=22)
    samples = M.perturbed_samples(base, 6, seed=7)
    _case(f"chi batch 6 samples n={n}")
    b = lt.ChiBatch(len(samples), base.type, n, frame, nimages, images)
    try:
        chi, found, status = b.run(samples)
        for s, L in enumerate(samples):
            st1, chi1 = lt.chi_barycentersource(L, n, frame, nimages, images)
            assert status[s] == st1
            assert chi[s] == chi1, s
            ost, ochi = oracle.chi_barycentersource(L, n, frame, nimages, images)
            assert ost == status[s]
            assert abs(chi[s] - ochi) <= 1e-9 * abs(ochi), s
            # delensed images of this sample vs the oracle's on the oracle gradient
            pos, loc = b.images(s)
            src = oracle.source_barycenters(nimages, images, L)
            gx, gy = oracle.gradient_grid(L, M.Grid(frame, n, n))
            _, opos, oloc, _ = oracle.delense_barycenter(len(nimages), n, frame, src, gx.ravel(), gy.ravel())
            np.testing.assert_array_equal(loc, oloc)
            np.testing.assert_allclose(pos, opos, rtol=0, atol=1e-12)
        # a second run with the samples in reverse order gives the reversed results
        chi_r, found_r, status_r = b.run(samples[::-1])
        np.testing.assert_array_equal(chi_r, chi[::-1])
        np.testing.assert_array_equal(found_r, found[::-1])
    finally:
        b.close()


def test_chi_rejects_bad_families(lt):
    lens = M.group_lens(seed=5)
    frame = M.grid_param(-100.0, 100.0, -100.0, 100.0)
    images = M.galaxies_array(3)
    with pytest.raises(lt.LtError):
        lt.ChiBatch(1, lens.type, 32, frame, np.array([0, 3], dtype=np.int32), images)
    with pytest.raises(lt.LtError):
        lt.ChiBatch(1, lens.type, 32, frame, np.array([21], dtype=np.int32), M.galaxies_array(21))
