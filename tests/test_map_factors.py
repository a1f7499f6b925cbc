"""Scale factors of the reference map modes (lt_map_factors; host only, no GPU).

Pinned by physics with independent constants (CODATA c, G, the solar mass, the parsec): the mass
map in 1e12 Msun/kpc^2 (mode 4) is the critical surface density c^2 D_s / (4 pi G D_l D_ls), in
g/cm^2 (mode 2) the same in other units, per pixel (mode 3) times the pixel area at the lens; the
ratio r is D_ls / D_s with distances from scipy quadrature. The library's constants are the
reference's (src/structure_hpc.hpp:63-68), rounded there to about 1e-4."""
import math

import pytest
from scipy.integrate import quad

import ltmodel as M

C_KMS = 299792.458
G_MPC = 4.300917e-9          # Mpc (km/s)^2 / Msun
MSUN_G = 1.98847e33
MPC_CM = 3.0856776e24


@pytest.fixture(scope="module")
def lt():
    import lthip
    return lthip


def _dm(z1, z2, om):
    return quad(lambda t: 1 / math.sqrt(om * (1 + t) ** 3 + 1 - om), z1, z2, epsabs=1e-14, epsrel=1e-12)[0]


def test_factors_against_physics(lt):
    H0, om, zl, zs = 70.0, 0.3, 0.35, 1.8
    cosmo = M.flat_lcdm(H0, om)
    dh = C_KMS / H0                                  # Mpc
    dl = dh * _dm(0, zl, om) / (1 + zl)
    ds = dh * _dm(0, zs, om) / (1 + zs)
    dls = dh * _dm(zl, zs, om) / (1 + zs)
    sigma_crit = C_KMS ** 2 / (4 * math.pi * G_MPC) * ds / (dl * dls)   # Msun / Mpc^2
    r, m1 = lt.map_factors(lt.MAP_AMPLI, 1, cosmo, zl, zs)
    assert r == pytest.approx(dls / ds, rel=1e-9) and m1 == 1.0
    _, m4 = lt.map_factors(lt.MAP_MASS, 4, cosmo, zl, zs)
    assert m4 == pytest.approx(sigma_crit / 1e6 / 1e12, rel=3e-4)
    _, m2 = lt.map_factors(lt.MAP_MASS, 2, cosmo, zl, zs)
    assert m2 == pytest.approx(sigma_crit * MSUN_G / MPC_CM ** 2, rel=3e-4)
    dx, dy = 0.2, 0.3
    _, m3 = lt.map_factors(lt.MAP_MASS, 3, cosmo, zl, zs, dx, dy)
    kpc_per_arcsec = dl * 1e3 * math.pi / 648000
    assert m3 == pytest.approx(m4 * kpc_per_arcsec ** 2 * dx * dy, rel=3e-4)
    _, m1m = lt.map_factors(lt.MAP_MASS, 1, cosmo, zl, zs)
    assert m1m == 1.0
    # potential: mode 1 scales by D_ls / D_s, mode 2 by c^2 / (2 pi G) / D_l in g/cm^2 units
    _, p1 = lt.map_factors(lt.MAP_POT, 1, cosmo, zl, zs)
    assert p1 == r
    _, p2 = lt.map_factors(lt.MAP_POT, 2, cosmo, zl, zs)
    assert p2 == pytest.approx(C_KMS ** 2 / (2 * math.pi * G_MPC) / dl * MSUN_G / MPC_CM ** 2, rel=3e-4)


def test_factors_errors(lt):
    cosmo = M.flat_lcdm()
    with pytest.raises(lt.LtError):
        lt.map_factors(lt.MAP_MASS, 9, cosmo, 0.3, 1.0)
    with pytest.raises(lt.LtError):
        lt.map_factors(99, 1, cosmo, 0.3, 1.0)
