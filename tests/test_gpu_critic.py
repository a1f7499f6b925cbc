"""Critical lines (marching squares on the GPU's det A plane) and caustics.

The device contour must equal the oracle's contour of the same planes bit for bit (segments, kinds,
count, order), on the real A2744 model with a pitched plane; count-only and overflow calls must
agree with it. lt_critic_lines is pinned by closed forms: for a singular isothermal sphere the
critical line is the circle r = b0 D_ls/D_s and the caustic its centre (both within the linear
interpolation's O(h^2)); a cored dPIE has an inner radial and an outer tangential line."""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest

import ltmodel as M
import pyoracle as O
from conftest import GOLDEN

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    import lthip
    lthip.check(lthip.lib().lt_init(0))
    return lthip


def _fields_device(hip, lens, grid, r, pitch, precision=0):
    m = hip.Model(lens, hip.MAP_HESSIAN)
    kap = hip.DeviceArray(pitch * grid.ny)
    det = hip.DeviceArray(pitch * grid.ny)
    kap.fill_bytes(0)
    det.fill_bytes(0)
    m.map(hip.MAP_FIELDS, 0, [kap.ptr, 0, 0, det.ptr, 0], grid, r=r, precision=precision, pitch=pitch)
    hip.sync()
    m.close()
    return kap, det


def test_device_contour_equals_oracle(hip):
    d = hip.load_par(os.path.join(GOLDEN, "par", "A2744_SLspec.par"))
    lens = d["lens"]
    grid = M.Grid(M.grid_param(-70.0, 70.0, -60.0, 50.0), 300, 240)
    r = hip.cosmo_dlsds(0.308, 2.0, d["cosmo"])
    pitch = grid.nx + 7
    print(f"[case] A2744 critical lines: {lens.n} halos, {grid.nx}x{grid.ny}, pitch {pitch}", flush=True)
    kap, det = _fields_device(hip, lens, grid, r, pitch)
    try:
        hk = kap.download().reshape(grid.ny, pitch)[:, :grid.nx]
        hd = det.download().reshape(grid.ny, pitch)[:, :grid.nx]
        want, wkinds = O.critic_lines(hd, hk, grid)
        assert len(want) > 200 and (wkinds == 0).any() and (wkinds == 1).any()
        got, kinds, total, st = hip.critic_device(det.ptr, kap.ptr, grid, 1 << 16, pitch=pitch)
        assert st == hip.LT_OK and total == len(want)
        np.testing.assert_array_equal(got, want)
        np.testing.assert_array_equal(kinds, wkinds)
        # count only; then room for 10: the first 10 and LT_ERR_OVERFLOW
        _, _, total0, st0 = hip.critic_device(det.ptr, 0, grid, 0, pitch=pitch)
        assert st0 == hip.LT_OK and total0 == total
        part, pk, total10, st10 = hip.critic_device(det.ptr, kap.ptr, grid, 10, pitch=pitch)
        assert st10 == hip.LT_ERR_OVERFLOW and total10 == total
        np.testing.assert_array_equal(part, want[:10])
        np.testing.assert_array_equal(pk, wkinds[:10])
    finally:
        kap.close()
        det.close()


def test_sis_circle_and_point_caustic(hip):
    cosmo = M.flat_lcdm()
    zl, zs = 0.3, 1.5
    lens = M.LensModel.from_halos([{"type": 5, "x": 0.0, "y": 0.0, "sigma": 900.0, "e": 0.0}], zl=zl)
    r = hip.cosmo_dlsds(zl, zs, cosmo)
    theta_e = lens.b0[0] * r
    n, half = 256, 2.5 * theta_e
    frame = M.grid_param(-half, half, -half, half)
    h = 2 * half / (n - 1)
    print(f"[case] SIS critical circle theta_E={theta_e:.4f} n={n}", flush=True)
    crit, caus, kinds, total = hip.critic_lines(lens, frame, n, n, cosmo, zl, zs)
    assert total == len(crit) > 100
    for px, py in ((crit[:, 0], crit[:, 1]), (crit[:, 2], crit[:, 3])):
        assert np.abs(np.hypot(px, py) - theta_e).max() < h * h / theta_e
    assert np.hypot(caus[:, 0], caus[:, 1]).max() < h * h / theta_e + 1e-12
    assert (kinds == 0).all()
    # closed and clockwise
    assert sorted(map(tuple, crit[:, :2])) == sorted(map(tuple, crit[:, 2:]))
    area = 0.5 * np.sum(crit[:, 0] * crit[:, 3] - crit[:, 2] * crit[:, 1])
    assert area < 0 and abs(-area / (math.pi * theta_e ** 2) - 1) < 1e-3


def test_cored_dpie_radial_and_tangential(hip):
    cosmo = M.flat_lcdm()
    zl, zs = 0.3, 2.0
    lens = M.LensModel.from_halos([{"type": 81, "x": 0.0, "y": 0.0, "sigma": 1100.0, "e": 0.2, "theta": 0.4,
                                    "rcore": 2.0, "rcut": 300.0}], zl=zl)
    frame = M.grid_param(-60.0, 60.0, -60.0, 60.0)
    print("[case] cored dPIE critical lines 400x400", flush=True)
    for prec in (0, 1):
        crit, caus, kinds, total = hip.critic_lines(lens, frame, 400, 400, cosmo, zl, zs, precision=prec)
        assert total == len(crit) > 100
        rad = np.hypot(crit[kinds == 1, 0], crit[kinds == 1, 1])
        tan = np.hypot(crit[kinds == 0, 0], crit[kinds == 0, 1])
        assert len(rad) > 10 and len(tan) > 10 and rad.max() < tan.min()
        # the tangential line maps to a small diamond-like caustic inside the radial line's caustic
        ct = np.hypot(caus[kinds == 0, 0], caus[kinds == 0, 1])
        cr = np.hypot(caus[kinds == 1, 0], caus[kinds == 1, 1])
        assert np.median(ct) < np.median(cr)


SIS_PAR = """runmode
\tampli 1 101 1.5 ampli_sis
\tend
source
\tz_source 1.5
\tend
cosmologie
\tH0 70.
\tomega 0.3
\tlambda 0.7
\tend
champ
\tdmax 30.
\tend
potentiel S
\tprofil 5
\tx_centre 0.
\ty_centre 0.
\tellipticite 0.
\tv_disp 700.
\tz_lens 0.3
\tend
cline
\tnplan 1 1.5
\tdmax 20.
\tnbgridcells 201
\tend
fini
"""


def test_maps_tool_critical_lines(hip, tmp_path):
    """lenstool_maps with a cline block: critic_0.dat holds the SIS critical circle r = b0 D_ls/D_s
    and caustic_0.dat points at the centre, both within the O(h^2) of the interpolation."""
    tool = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "lenstool-hpc_amd", "lenstool_maps")
    par = tmp_path / "sis.par"
    par.write_text(SIS_PAR)
    print("[case] lenstool_maps cline: SIS, 201x201 over +-20", flush=True)
    r = subprocess.run([tool, str(par), "--out", str(tmp_path), "--quiet"], capture_output=True, text=True, timeout=120)
    print(r.stdout[-1000:], r.stderr[-1000:], flush=True)
    assert r.returncode == 0
    crit = np.loadtxt(tmp_path / "critic_0.dat")
    caus = np.loadtxt(tmp_path / "caustic_0.dat")
    d = hip.load_par(str(par))
    theta_e = d["lens"].b0[0] * hip.cosmo_dlsds(0.3, 1.5, d["cosmo"])
    h = 40.0 / 200
    assert len(crit) > 50 and len(caus) == len(crit) and (crit[:, 4] == 0).all()
    # 10 decimals in the file
    assert np.abs(np.hypot(crit[:, 0], crit[:, 1]) - theta_e).max() < h * h / theta_e + 1e-9
    assert np.hypot(caus[:, 0], caus[:, 1]).max() < h * h / theta_e + 1e-9
    assert os.path.exists(tmp_path / "ampli_sis.fits")


def test_external_shear_lines_without_caustics(hip):
    """Type-14 halos have a Hessian but no deflection: the lines are the contour of the model's own
    det-A plane, the caustics are NaN and the call reports LT_ERR_TYPE."""
    cosmo = M.flat_lcdm()
    zl, zs = 0.3, 2.0
    lens = M.LensModel.from_halos([{"type": 81, "x": 0.0, "y": 0.0, "sigma": 1000.0, "e": 0.2, "theta": 0.4,
                                    "rcore": 2.0, "rcut": 300.0},
                                   {"type": 14, "x": 0.0, "y": 0.0, "b0": 0.0, "epot": 0.03, "theta": 0.7}], zl=zl)
    n = 300
    frame = M.grid_param(-60.0, 60.0, -60.0, 60.0)
    grid = M.Grid(frame, n, n)
    r = hip.cosmo_dlsds(zl, zs, cosmo)
    print("[case] dPIE + external shear critical lines 300x300", flush=True)
    st = hip.lib().lt_critic_lines(C.byref(cosmo), C.byref(frame), C.byref(lens.soa()), lens.n, zl, zs, n, n, 0,
                                   None, None, None, 0, C.byref(C.c_longlong(0)))
    assert st == hip.LT_OK  # no caustic array: nothing unavailable
    crit, caus, kinds, total = hip.critic_lines(lens, frame, n, n, cosmo, zl, zs)
    assert total == len(crit) > 100 and np.isnan(caus).all()
    assert b"caustics unavailable" in hip.lib().lt_last_error()
    kap, det = _fields_device(hip, lens, grid, r, n)
    try:
        want, wk, wt, wst = hip.critic_device(det.ptr, kap.ptr, grid, 1 << 16)
        assert wst == hip.LT_OK and wt == total
        np.testing.assert_array_equal(crit, want)
        np.testing.assert_array_equal(kinds, wk)
    finally:
        kap.close()
        det.close()
