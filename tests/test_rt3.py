"""Oracle pins from the reference's result test (Benchmarks/ResultTest, tests/rt3_data.py): what the
original Lenstool computed for a 6-halo PIEMD lens -- source positions of 30 images (the type-8
gradient, D_ls/D_s and the parameter-file reader together), their amplifications (the type-8
Hessian) and chi^2 = 0 of the model's own images (barycentres, delensing, matching) -- and the lens
equation itself for the source barycentres: images that solve theta - dr alpha(theta) = s to
rounding give back s."""
import numpy as np
import pytest

import ltmodel as M
import rt3_data as R


@pytest.fixture(scope="module")
def rt3():
    import lthip
    return R.load(lthip)


def test_sources_match_lenstool(oracle, rt3):
    d, img, src, _ = rt3
    L = d["lens"]
    assert d["info"]["nhalos"] == 6 and (L.type == 8).all()
    dr = oracle.cosmo_dlsds(R.Z_LENS, 1.5, d["cosmo"])
    for (fam, x, y), (sx, sy) in zip(img, src):
        gx, gy = oracle.point_gradient(L, x, y)
        # source.dat has 7 decimals
        assert abs(x - dr * gx - sx) <= 5.01e-8 and abs(y - dr * gy - sy) <= 5.01e-8, (fam, x, y)


def test_amplifications_match_lenstool(oracle, rt3):
    d, img, _, rows = rt3
    L = d["lens"]
    dr = oracle.cosmo_dlsds(R.Z_LENS, 1.5, d["cosmo"])
    for fam, x, y in img:
        a, b, c, _ = oracle.point_hessian(L, x, y)
        mu = 1.0 / abs((1 - dr * a) * (1 - dr * c) - (dr * b) ** 2)
        amp = R.amp_rows(rows, int(fam), x, y)
        assert len(amp) >= 1
        assert np.abs(amp - mu).min() <= R.amp_tolerance(mu), (fam, x, y, mu, amp)


def test_chi2_of_model_images_vanishes(oracle, rt3):
    """chires.dat: chi^2 = 0 for the model's own images. On the reference's grid path every image is
    found (no 1000-per-image penalty), each within a fraction of a cell, and chi^2 falls as the
    square of the cell size."""
    d, img, _, _ = rt3
    ims = d["images"].copy()
    assert len(ims) == 30 and d["nimages"].sum() == 30
    ims["dr"] = [oracle.cosmo_dlsds(R.Z_LENS, z, d["cosmo"]) for z in ims["redshift"]]
    chis = []
    for n in (500, 1000, 2000):
        st, chi = oracle.chi_barycentersource(d["lens"], n, d["frame"], d["nimages"], ims)
        cell = (d["frame"].xmax - d["frame"].xmin) / (n - 1)
        assert st == 0 and chi < len(ims) * (0.5 * cell) ** 2, (n, chi)
        chis.append(chi)
    assert chis[1] < chis[0] / 2 and chis[2] < chis[1] / 2


def _solve_images(oracle, L, dr, guesses, s):
    """Newton on theta - dr alpha(theta) = s from each guess (Jacobian 1 - dr H)."""
    out = []
    for x, y in guesses:
        for _ in range(40):
            gx, gy = oracle.point_gradient(L, x, y, extended=True)
            fx, fy = x - dr * gx - s[0], y - dr * gy - s[1]
            if max(abs(fx), abs(fy)) < 1e-14:
                break
            a, b, c, _ = oracle.point_hessian(L, x, y)
            j11, j12, j22 = 1 - dr * a, -dr * b, 1 - dr * c
            det = j11 * j22 - j12 * j12
            x, y = x - (j22 * fx - j12 * fy) / det, y - (j11 * fy - j12 * fx) / det
        out.append((x, y))
    return out


def test_source_barycenters_invert_lens_equation(oracle, rt3):
    """orc_source_barycenters (chi_CPU_barycenter.cpp:173-210): the mean over a family of
    theta - dr alpha(theta). For images solving the lens equation for s exactly (Newton, to 1e-14),
    every term is s, so the barycentre is s to rounding, whatever the family size; a wrong dr, sign
    or image index moves it by arcseconds."""
    d, img, src, _ = rt3
    L = d["lens"]
    fams = img[:, 0].astype(int)
    nimages, images, want = [], [], []
    for k, f in enumerate(np.unique(fams)):
        sel = np.flatnonzero(fams == f)
        z = 1.0 + 0.5 * k  # a different source plane per family: dr from each family's first image
        dr = oracle.cosmo_dlsds(R.Z_LENS, z, d["cosmo"])
        s = src[sel[0]] * (0.9 + 0.02 * k)
        sol = _solve_images(oracle, L, dr, img[sel, 1:3], s)
        nimages.append(len(sol))
        for x, y in sol:
            images.append((x, y, z, dr))
        want.append(s)
    gal = M.galaxies_array(len(images))
    gal["cx"], gal["cy"], gal["redshift"], gal["dr"] = np.array(images).T
    got = oracle.source_barycenters(np.array(nimages, np.int32), gal, L)
    np.testing.assert_allclose(np.c_[got["cx"], got["cy"]], np.array(want), rtol=0, atol=1e-10)
    # the images' own dr is the one used: with dr halved, the barycentres move far away
    gal["dr"] *= 0.5
    off = oracle.source_barycenters(np.array(nimages, np.int32), gal, L)
    assert np.hypot(off["cx"] - np.array(want)[:, 0], off["cy"] - np.array(want)[:, 1]).min() > 0.1
