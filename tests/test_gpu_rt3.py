"""The reference's result test (Benchmarks/ResultTest, tests/rt3_data.py) through the HIP library: the
6-halo PIEMD model read by lt_par; the FP64 point gradients at the 30 images give Lenstool's source
positions (source.dat, 7 decimals), the FP64 Hessian maps at the images its amplifications (dist.dat,
3 decimals), and the chi^2 of the model's own images on the GPU's gradient grid + delensing +
matching is the oracle's and vanishes with the cell size (chires.dat: chi^2 = 0)."""
import numpy as np
import pytest

import ltmodel as M
import rt3_data as R
from parity import EXT_SHARE, REL

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    import lthip
    lthip.check(lthip.lib().lt_init(0))
    return lthip


@pytest.fixture(scope="module")
def rt3(hip):
    return R.load(hip)


def test_sources_match_lenstool(hip, oracle, rt3):
    d, img, src, _ = rt3
    L = d["lens"]
    dr = hip.cosmo_dlsds(R.Z_LENS, 1.5, d["cosmo"])
    print(f"[case] RT-3 sources: {L.n} PIEMD halos, {len(img)} images", flush=True)
    gx, gy = hip.point_gradients(L, img[:, 1], img[:, 2])
    np.testing.assert_allclose(img[:, 1] - dr * gx, src[:, 0], rtol=0, atol=5.01e-8)
    np.testing.assert_allclose(img[:, 2] - dr * gy, src[:, 1], rtol=0, atol=5.01e-8)
    for k, (_, x, y) in enumerate(img):
        ex, ey = oracle.point_gradient(L, x, y, extended=True)
        ax, ay = oracle.point_gradient_abs(L, x, y)
        cg, _ = oracle.point_cond(L, x, y)
        assert abs(gx[k] - ex) <= REL * ax + EXT_SHARE * cg and abs(gy[k] - ey) <= REL * ay + EXT_SHARE * cg


def test_amplifications_match_lenstool(hip, oracle, rt3):
    d, img, _, rows = rt3
    L = d["lens"]
    dr = hip.cosmo_dlsds(R.Z_LENS, 1.5, d["cosmo"])
    print(f"[case] RT-3 amplifications at {len(img)} images (1x1 Hessian grids)", flush=True)
    for fam, x, y in img:
        g = M.Grid(M.grid_param(x, x, y, y), 1, 1, step=(1.0, 1.0))
        a, b, c, _ = (p[0, 0] for p in hip.hessian_grid(L, g))
        mu = 1.0 / abs((1 - dr * a) * (1 - dr * c) - (dr * b) ** 2)
        amp = R.amp_rows(rows, int(fam), x, y)
        assert len(amp) >= 1 and np.abs(amp - mu).min() <= R.amp_tolerance(mu), (fam, x, y, mu, amp)


def test_chi2_of_model_images(hip, oracle, rt3):
    d, _, _, _ = rt3
    ims = d["images"].copy()
    ims["dr"] = [hip.cosmo_dlsds(R.Z_LENS, z, d["cosmo"]) for z in ims["redshift"]]
    chis = []
    for n in (500, 1000, 2000):
        print(f"[case] RT-3 chi^2 on a {n}^2 delensing grid", flush=True)
        st, chi = hip.chi_barycentersource(d["lens"], n, d["frame"], d["nimages"], ims)
        ost, ochi = oracle.chi_barycentersource(d["lens"], n, d["frame"], d["nimages"], ims)
        cell = (d["frame"].xmax - d["frame"].xmin) / (n - 1)
        assert st == 0 and ost == 0
        assert abs(chi - ochi) <= 1e-9 * max(ochi, 1e-12)
        assert chi < len(ims) * (0.5 * cell) ** 2
        chis.append(chi)
    assert chis[1] < chis[0] / 2 and chis[2] < chis[1] / 2
