"""The Lenstool parameter-file reader (lt_par_*; runs on the CPU, no GPU needed).

Pinned by the halo table the reference recorded after reading A2744_SLspec.par (its debug print of
all 258 SOA halos, 6 decimals: tests/golden/par/a2744_recorded_halos.npz), and on m1931.par by the
closed forms of the conversions the reference applies (b0 = 6 pi_c2 sigma^2, epot from the mass
ellipticity, kpc -> arcsec with d0 / h D_A(z), the potfile scaling relations), with D_A from scipy
quadrature rather than the library's own integrator."""
import math
import os

import numpy as np
import pytest
from scipy.integrate import quad

import ltmodel as M
from conftest import GOLDEN

PAR = os.path.join(GOLDEN, "par")
PI_C2 = 7.209970e-06
D0 = 29.068701


@pytest.fixture(scope="module")
def lt():
    import lthip
    return lthip


def _da_flat(z, om):
    """Angular diameter distance in units of c/H0, flat LambdaCDM."""
    return quad(lambda t: 1 / math.sqrt(om * (1 + t) ** 3 + 1 - om), 0, z, epsabs=1e-15, epsrel=1e-13)[0] / (1 + z)


def test_a2744_matches_recorded_table(lt):
    d = lt.load_par(os.path.join(PAR, "A2744_SLspec.par"))
    rec = np.load(os.path.join(PAR, "a2744_recorded_halos.npz"))
    L, e = d["lens"], d["extra"]
    assert d["info"]["nhalos"] == 258 and d["info"]["nhalos_main"] == 6 and d["info"]["npotfiles"] == 3
    got = {"x": L.x, "y": L.y, "b0": L.b0, "vdisp": e["vdisp"], "ellipticity": e["ellipticity"], "epot": L.epot,
           "angle": L.angle, "rcut": L.rcut, "z": L.z}
    for k, v in got.items():
        # the record has 6 decimals
        np.testing.assert_allclose(v, rec[k], rtol=0, atol=5.01e-7, err_msg=k)
    np.testing.assert_array_equal(L.type, rec["type"].astype(np.int32))
    # rcore: equal except the six clumps of the two potfiles with corekpc 150 and 400 kpc, which the
    # recorded run read as zero (its table prints "corekpc = 0" for them): this reader applies the
    # scaling relation to them as the reference's current reader does
    core_clumps = np.array([6, 7, 8, 255, 256, 257])
    other = np.setdiff1d(np.arange(258), core_clumps)
    np.testing.assert_allclose(L.rcore[other], rec["rcore"][other], rtol=0, atol=5.01e-7)
    assert (L.rcore[core_clumps] > 100).all() and (rec["rcore"][core_clumps] == 0).all()
    # cos / sin of the angles
    np.testing.assert_allclose(e["anglecos"], np.cos(L.angle), rtol=0, atol=1e-15)
    # images: 83 in 29 families, from the multfile of the image block
    assert d["info"]["nimagestot"] == 83 and d["info"]["nsets"] == 29 and d["nimages"].sum() == 83


def test_m1931_closed_forms(lt):
    d = lt.load_par(os.path.join(PAR, "m1931.par"))
    info, L, e, c, f = d["info"], d["lens"], d["extra"], d["cosmo"], d["frame"]
    assert (c.H0, c.h, c.omegaM, c.omegaX) == (70.0, 1.4, 0.3, 0.7)
    assert (f.xmin, f.xmax, f.ymin, f.ymax) == (-136.5, 136.5, -136.5, 136.5)
    assert info["nbgridcells"] == 4000 and info["z_source"] == 1.25
    assert (info["ampli"], info["ampli_cells"]) == (5, 4200)
    ncat = sum(1 for s in open(os.path.join(PAR, "m1931.cat")) if s.strip() and not s.startswith("#"))
    assert info["nhalos_main"] == 3 and info["nhalos"] == 3 + ncat
    d1 = D0 / 1.4 * _da_flat(0.35, 0.3)  # kpc per arcsec at z_lens
    # halo A: sigma 1000, e 0.3, 55 deg, rcore 50 kpc, rcut 1000 kpc
    assert L.b0[0] == pytest.approx(6 * PI_C2 * 1000.0 ** 2, rel=1e-15)
    assert L.epot[0] == pytest.approx((1 - math.sqrt(1 - 0.09)) / 0.3, rel=1e-15)
    assert L.angle[0] == pytest.approx(55.0 * math.pi / 180, rel=1e-15)
    assert L.rcore[0] == pytest.approx(50.0 / d1, rel=1e-12)
    assert L.rcut[0] == pytest.approx(1000.0 / d1, rel=1e-12)
    # halo B at (-4.03, -40.76); BCG: rcore 0.115 kpc, rcut 100 kpc, sigma 100
    assert (L.x[1], L.y[1]) == (-4.03, -40.76)
    assert L.rcore[2] == pytest.approx(0.115 / d1, rel=1e-12) and L.b0[2] == pytest.approx(6 * PI_C2 * 1e4, rel=1e-15)
    # first potfile clump: "853.0 -18.176113 -26.415253 0.51372004 0.34239 78.61 20.1248 0.3442"
    k = 3
    x, y = -18.176113, -26.415253
    dec = -26.575718 * math.pi / 180
    xr = (x / (-3600 * math.cos(dec)) + 292.95682 - 292.95682) * (-3600 * math.cos(dec))
    assert L.x[k] == pytest.approx(xr, abs=1e-9) and L.y[k] == pytest.approx(y, abs=1e-9)
    a, b, mag = 0.51372004, 0.34239, 20.1248
    ell = (a * a - b * b) / (a * a + b * b)
    assert e["ellipticity"][k] == pytest.approx(ell, rel=1e-14)
    assert L.epot[k] == pytest.approx((1 - math.sqrt(1 - ell * ell)) / ell, rel=1e-14)
    f10 = 10 ** (0.4 * (19.65 - mag))
    assert e["vdisp"][k] == pytest.approx(158.0 * f10 ** (1 / 4.0), rel=1e-13)
    assert L.rcut[k] == pytest.approx(10.0 / d1 * f10 ** (2 / 4.0), rel=1e-12)
    assert L.rcore[k] == pytest.approx(0.15 / d1 * f10 ** 0.5, rel=1e-12)
    assert L.angle[k] == pytest.approx(78.61 * math.pi / 180, rel=1e-15)
    # dlsds for the source plane of the run
    ds = _da_flat(1.25, 0.3)
    dls = (quad(lambda t: 1 / math.sqrt(0.3 * (1 + t) ** 3 + 0.7), 0.35, 1.25, epsabs=1e-15, epsrel=1e-13)[0]
           / (1 + 1.25))
    assert L.dlsds[0] == pytest.approx(dls / ds, rel=1e-10)
    # images: families by the integer part of the id, 10 families, 28 images
    assert info["nsets"] == 10 and info["nimagestot"] == 28
    assert d["images"]["cx"][0] == 292.96063 and d["images"]["redshift"][0] == 1.834


def test_missing_files_are_errors(lt, tmp_path):
    with pytest.raises(lt.LtError):
        lt.load_par(str(tmp_path / "absent.par"))
    p = tmp_path / "bad.par"
    p.write_text("runmode\n\tnbgridcells 10\n\tend\npotfile\n\tfilein 1 nowhere.cat\n\ttype 81\n\tend\nfini\n")
    with pytest.raises(lt.LtError):
        lt.load_par(str(p))
    q = tmp_path / "noz.par"
    q.write_text("potentiel 1\n\tprofil 81\n\tv_disp 300.\n\tend\nfini\n")
    with pytest.raises(lt.LtError):
        lt.load_par(str(q))
