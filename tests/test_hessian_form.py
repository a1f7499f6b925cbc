"""The Hessian kernels (lenstool-hpc_amd/csrc/lt_terms64.h hess_piemd, lt_kernels32.hip hess_piemd32)
evaluate the I_{0.5} Hessian of a core radius r in scaled halo-frame coordinates X = cx1 a x,
Y = cx1 b y, W = sqrt(cx1^2 r^2 + X^2 + Y^2), Q = W + r + i k1 X, U = Y / W + k5, as
    xx + yy = |Q|^2 / W,   xx - yy = -(X^2 / W + U (k3 W - k4 Y)),   xy = k1 X - k2 X U,
each over |Q|^2 and times 2 sqrt(e) a^2 (b0 t05 ci), instead of the reference's quotients of complex
numbers (src/gradient2.cpp:54-254). This pins that form: the formula, evaluated at 40 digits with the
constants the library packs (lenstool-hpc_amd/csrc/lt_lens.cpp), equals the oracle's extended-precision
restatement of the reference's form, at points over the field, near a centre and far outside the cut
radius, for PIEMD and dPIE halos of random ellipticity, angle and radii. It also checks the trace
identity (xx + yy) / |Q|^2 = 1 / W on its own, against the numerators of the form without it."""
import mpmath as mp
import numpy as np

import ltmodel as M
import pyoracle as O

mp.mp.dps = 40


def _terms(eps, r, X, Y):
    """Per radius: (xx + yy, xx - yy, xy) over |Q|^2, with the trace from 1 / W; and the trace from
    the numerators xx = Re Q - X^2 / W, yy = Re Q + U (k3 W - k4 Y)."""
    s = mp.sqrt(eps)
    a, b, cx1 = 1 / (1 + eps), 1 / (1 - eps), (1 - eps) / (1 + eps)
    k1, k2, k3, k4, k5 = 2 * s * a / cx1, b / a, 2 * s * b / cx1, 1 / cx1 ** 2, 2 * s * a
    W = mp.sqrt(cx1 ** 2 * r * r + X * X + Y * Y)
    Qr, Qi = W + r, k1 * X
    U = Y / W + k5
    q2 = Qr * Qr + Qi * Qi
    dd = -(X * X / W + U * (k3 * W - k4 * Y))
    xy = Qi - k2 * X * U
    xx, yy = Qr - X * X / W, Qr + U * (k3 * W - k4 * Y)
    return 1 / W, dd / q2, xy / q2, (xx + yy) / q2


def kernel_form(t, x0, y0, b0, cs, sn, eps, rc, rcut, px, py):
    eps, rc, rcut = mp.mpf(eps), mp.mpf(rc), mp.mpf(rcut)
    cs, sn = mp.mpf(cs), mp.mpf(sn)
    tx, ty = mp.mpf(px) - mp.mpf(x0), mp.mpf(py) - mp.mpf(y0)
    a, b, cx1 = 1 / (1 + eps), 1 / (1 - eps), (1 - eps) / (1 + eps)
    X = cx1 * a * (tx * cs + ty * sn)
    Y = cx1 * b * (ty * cs - tx * sn)
    Ta, Da, Ba, Ta2 = _terms(eps, rc, X, Y)
    if t == 81:
        Tb, Db, Bb, Tb2 = _terms(eps, rcut, X, Y)
        dT, dD, dB, dT2 = Ta - Tb, Da - Db, Ba - Bb, Ta2 - Tb2
    else:
        dT, dD, dB, dT2 = Ta, Da, Ba, Ta2
    s = mp.sqrt(eps)
    t05 = rcut / (rcut - rc) if t == 81 else 1
    amph = mp.mpf(b0) * t05 * (1 - eps * eps) / (2 * s)
    e2a2 = 2 * s * a * a
    cos2, sin2 = cs * cs - sn * sn, 2 * cs * sn
    T = amph * e2a2 * dT
    D = amph * e2a2 * (cos2 * dD - 2 * sin2 * dB)
    B = amph * e2a2 * (sin2 * dD / 2 + cos2 * dB)
    return (T + D) / 2, B, (T - D) / 2, abs(dT - dT2) / abs(dT2)


def test_hessian_form_equals_reference_form():
    rng = np.random.default_rng(77)
    worst, worst_trace = 0.0, 0.0
    for trial in range(16):
        t = (8, 81)[trial % 2]
        rc = float(10 ** rng.uniform(-1.5, 0.7))
        halo = dict(type=t, x=rng.uniform(-3, 3), y=rng.uniform(-3, 3), sigma=rng.uniform(150, 900),
                    e=rng.uniform(0.05, 0.6), theta=rng.uniform(0, np.pi), rcore=rc,
                    rcut=rc * float(10 ** rng.uniform(0.5, 2.5)))
        lens = M.LensModel.from_halos([halo])
        h = (lens.x[0], lens.y[0], lens.b0[0], np.cos(lens.angle[0]), np.sin(lens.angle[0]), lens.epot[0],
             lens.rcore[0], lens.rcut[0])
        far = 30 * lens.rcut[0]
        for lo_x, hi_x, lo_y, hi_y in [(-60, 60, -50, 50),
                                       (h[0] + 0.5 * rc, h[0] + 3 * rc, h[1] - 2 * rc, h[1] + rc),
                                       (far, 1.2 * far, 0.3 * far, 0.5 * far)]:
            for px in np.linspace(lo_x, hi_x, 4):
                for py in np.linspace(lo_y, hi_y, 4):
                    ka, kb, kc, tr = kernel_form(t, *h, px, py)
                    ea, eb, ec, _ = O.point_hessian(lens, px, py, extended=True)
                    sa, sb, sc, _ = O.point_hessian_abs(lens, px, py)
                    scale = sa + sb + sc
                    err = max(abs(float(ka) - ea), abs(float(kb) - eb), abs(float(kc) - ec))
                    worst = max(worst, err / scale)
                    worst_trace = max(worst_trace, float(tr))
    print(f"max difference / sum of |a|, |b|, |c|, kernel form (40 digits) vs reference form (extended): "
          f"{worst:.2e}; trace identity: {worst_trace:.2e}")
    assert worst < 1e-15
    assert worst_trace < 1e-30
