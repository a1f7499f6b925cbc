"""The potential kernels evaluate the I_{0.5} potential of a core radius r as
    phi_r = x arg Q + y ln(S / |Q|) - 2 sqrt(e) r ln((1 + e) |Q| / (2 r))          (halo frame),
Q = r + cx1 w + 2i sqrt(e) a^2 x, S = w + cx1 r + 2 sqrt(e) a b y, w = sqrt(r^2 + a^2 x^2 + b^2 y^2),
a = 1/(1+e), b = 1/(1-e), cx1 = (1-e)/(1+e) (a dPIE: the core radius minus the cut radius, both times
the amplitude b0 t05 (1 - e^2) / (2 sqrt(e))), instead of the reference's complex cosh / sinh / log of
eta and zeta (src/potential_GPU.cu:47-72). This pins that identity: the formula, evaluated at 40
digits, equals the oracle's extended-precision restatement of the reference's form (itself pinned by
finite differences against the gradient in test_oracle.py) at points over the field, near a centre
and far outside the cut radius, for PIEMD and dPIE halos of random ellipticity and radii (the dPIE is
the difference of two such terms, so a PIEMD whose core radius is the cut radius checks that term
next to the centre)."""
import mpmath as mp
import numpy as np

import ltmodel as M
import pyoracle as O

mp.mp.dps = 40


def closed_form(t, x0, y0, b0, ang, eps, rc, rcut, px, py):
    eps, rc, rcut = mp.mpf(eps), mp.mpf(rc), mp.mpf(rcut)
    cs, sn = mp.cos(mp.mpf(ang)), mp.sin(mp.mpf(ang))
    tx, ty = mp.mpf(px) - mp.mpf(x0), mp.mpf(py) - mp.mpf(y0)
    x, y = tx * cs + ty * sn, ty * cs - tx * sn
    s = mp.sqrt(eps)
    a, b, q = 1 / (1 + eps), 1 / (1 - eps), (1 - eps) / (1 + eps)

    def one(r):
        w = mp.sqrt(r * r + a * a * x * x + b * b * y * y)
        Q = mp.mpc(r + q * w, 2 * s * a * a * x)
        S = w + r * q + 2 * s * a * b * y
        return x * mp.atan2(Q.imag, Q.real) + y * mp.log(S / abs(Q)) - 2 * s * r * mp.log((1 + eps) * abs(Q) / (2 * r))

    t05 = rcut / (rcut - rc) if t == 81 else 1
    amp = mp.mpf(b0) * t05 * (1 - eps * eps) / (2 * s)
    return amp * (one(rc) - (one(rcut) if t == 81 else 0))


def test_potential_closed_form_equals_reference_form():
    rng = np.random.default_rng(2024)
    worst = 0.0
    for trial in range(18):
        t = (8, 81)[trial % 2]
        rc = float(10 ** rng.uniform(-1.5, 0.7)) if trial < 16 else float(10 ** rng.uniform(0.5, 1.5))
        halo = dict(type=t, x=rng.uniform(-3, 3), y=rng.uniform(-3, 3), sigma=rng.uniform(150, 900),
                    e=rng.uniform(0.05, 0.6), theta=rng.uniform(0, np.pi), rcore=rc,
                    rcut=rc * float(10 ** rng.uniform(0.5, 2.5)))
        lens = M.LensModel.from_halos([halo])
        h = lens.x[0], lens.y[0], lens.b0[0], lens.angle[0], lens.epot[0], lens.rcore[0], lens.rcut[0]
        # a coarse field, a patch next to the centre (PIEMD: 0.01-0.3 core radii; dPIE: 0.5-3, since the
        # reference's form of the cut radius' term cancels as (rcut / d)^2 next to the centre, and at
        # 0.01 core radii even the extended oracle keeps only ~5e-14 there) and a patch far outside rcut
        far = 30 * lens.rcut[0]
        near = (0.01, 0.3, -0.2, 0.1) if t == 8 else (0.5, 3.0, -2.0, 1.0)
        for lo_x, hi_x, lo_y, hi_y in [(-60, 60, -50, 50),
                                       (h[0] + near[0] * rc, h[0] + near[1] * rc, h[1] + near[2] * rc,
                                        h[1] + near[3] * rc),
                                       (far, 1.2 * far, 0.3 * far, 0.5 * far)]:
            n = 5
            grid = M.Grid(M.grid_param(lo_x, hi_x, lo_y, hi_y), n, n)
            ext = O.potential_grid(lens, grid, extended=True)
            for j in range(n):
                for i in range(n):
                    v = closed_form(t, *h, grid.frame.xmin + i * grid.dx, grid.frame.ymin + j * grid.dy)
                    worst = max(worst, abs(float(v) - ext[j, i]) / abs(ext[j, i]))
    print(f"max relative difference, closed form (40 digits) vs reference form (extended): {worst:.2e}")
    assert worst < 2e-14
