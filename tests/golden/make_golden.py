"""Write golden vectors from the reference's own CPU functions, compiled from the reference sources
by `make -C oracle ref` (oracle/_ref/libltref.so: src/gradient2.cpp, src/grid_gradient2_CPU.cpp,
src/delense_CPU.cpp, src/chi_computation.cpp). Run once where the reference is mounted:

    make -C oracle ref && python tests/golden/make_golden.py

Outputs (numpy .npz, no pickles) next to this script:
  hess81_grid.npz    dPIE Hessian grid (gradient2_grid_CPU, dx/dy overload, istart/jstart offsets)
  hess81_points.npz  dPIE Hessian at near-core and far points (totalGradient2_SOA)
  delense.npz        delense_barycenter found images for fixed gradient planes and sources
  chi.npz            chi_computation on the delensed images with perturbed observed images
The lens models are synthetic (seeded) and stored in the files, so the tests need only numpy.
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "lenstool-hpc_amd"))

import ltmodel as M  # noqa: E402
import pyoracle as O  # noqa: E402


def lens_arrays(L: M.LensModel):
    return dict(type=L.type, x=L.x, y=L.y, b0=L.b0, angle=L.angle, epot=L.epot, rcore=L.rcore,
                rcut=L.rcut)


def hess_lens(seed=11):
    rng = np.random.default_rng(seed)
    halos = []
    for k in range(12):
        e = float(rng.uniform(0, 0.7))
        halos.append(dict(type=81, x=float(rng.uniform(-20, 20)), y=float(rng.uniform(-20, 20)),
                          sigma=float(rng.uniform(100, 900)), e=e, theta=float(rng.uniform(0, np.pi)),
                          rcore=float(10 ** rng.uniform(-1.3, 1)), rcut=float(10 ** rng.uniform(1, 3))))
    halos[3]["e"] = 0.0          # epot = 1e-5 (the reference's floor for e = 0)
    halos[5]["rcore"] = 0.02     # very small core
    return M.LensModel.from_halos(halos)


def main():
    R = O.ref_lib()
    if R is None:
        raise SystemExit("oracle/_ref/libltref.so missing: run `make -C oracle ref`")
    R.ref_set_threads(1)

    # 1. Hessian grid ----------------------------------------------------------------------------
    L = hess_lens()
    frame = M.grid_param(-25.0, 25.0, -22.0, 22.0)
    nx, ny, i0, j0 = 48, 40, 3, 2
    dx, dy = 50.0 / (nx + 2 * i0 - 1), 44.0 / (ny + 2 * j0 - 1)
    out = (M.LtMatrix * (nx * ny))()
    R.ref_gradient2_grid(out, C.byref(frame), C.byref(L.soa()), L.n, dx, dy, nx, ny, i0, j0)
    m = np.ctypeslib.as_array(out).view(np.float64).reshape(ny, nx, 4)
    np.savez(os.path.join(HERE, "hess81_grid.npz"), **lens_arrays(L),
             frame=np.array([frame.xmin, frame.xmax, frame.ymin, frame.ymax]),
             shape=np.array([nx, ny, i0, j0]), dxdy=np.array([dx, dy]),
             a=m[..., 0].copy(), b=m[..., 1].copy(), c=m[..., 2].copy(), d=m[..., 3].copy())

    # 2. Hessian at points near the cores (0.001 .. 3 rcore) and far away ------------------------
    rng = np.random.default_rng(12)
    pts = []
    for i in range(L.n):
        for f in (1e-3, 0.1, 0.7, 3.0):
            ang = rng.uniform(0, 2 * np.pi)
            pts.append((L.x[i] + f * L.rcore[i] * np.cos(ang), L.y[i] + f * L.rcore[i] * np.sin(ang)))
    pts += [(300.0, -250.0), (-1000.0, 40.0)]
    pts = np.array(pts)
    H = np.empty((len(pts), 4))
    for k, (px, py) in enumerate(pts):
        o = M.LtMatrix()
        p = M.LtPoint(px, py)
        R.ref_point_gradient2(C.byref(o), C.byref(p), C.byref(L.soa()), L.n)
        H[k] = (o.a, o.b, o.c, o.d)
    np.savez(os.path.join(HERE, "hess81_points.npz"), **lens_arrays(L), points=pts, hessian=H)

    # 3. delensing -------------------------------------------------------------------------------
    lens1 = M.group_lens(seed=5)
    n = 64
    dframe = M.grid_param(-100.0, 100.0, -100.0, 100.0)
    grid = M.Grid(dframe, n, n)
    gx, gy = O.gradient_grid(lens1, grid)
    gx = np.ascontiguousarray(gx.ravel())
    gy = np.ascontiguousarray(gy.ravel())
    nsets = 16
    src = M.galaxies_array(nsets)
    src["cx"] = rng.uniform(-6, 6, nsets)
    src["cy"] = rng.uniform(-6, 6, nsets)
    src["dr"] = rng.uniform(0.8, 0.95, nsets)
    pos = (M.LtPoint * (nsets * M.MAXIMPERSOURCE))()
    loc = (C.c_int * nsets)()
    numimg = C.c_int(0)
    R.ref_delense_barycenter(pos, loc, C.byref(numimg), nsets, n, C.byref(dframe), M.galaxy_ptr(src),
                             M.dptr(gx), M.dptr(gy))
    posa = np.ctypeslib.as_array(pos).view(np.float64).reshape(nsets, M.MAXIMPERSOURCE, 2).copy()
    loca = np.ctypeslib.as_array(loc).copy()
    assert loca.max() <= M.MAXIMPERSOURCE
    np.savez(os.path.join(HERE, "delense.npz"), n=np.array([n]),
             frame=np.array([dframe.xmin, dframe.xmax, dframe.ymin, dframe.ymax]), gx=gx, gy=gy,
             sources=src, image_pos=posa, locimagesfound=loca, numimg=np.array([numimg.value]))

    # 4. chi^2 on the delensed images, observed images = found ones + noise, families of 3 -------
    nimages = np.full(nsets, 3, dtype=np.int32)
    images = M.galaxies_array(3 * nsets)
    for s in range(nsets):
        for k in range(3):
            if k < loca[s]:
                images[3 * s + k]["cx"] = posa[s, k, 0] + rng.normal(0, 0.3)
                images[3 * s + k]["cy"] = posa[s, k, 1] + rng.normal(0, 0.3)
            else:
                images[3 * s + k]["cx"] = rng.uniform(-80, 80)
                images[3 * s + k]["cy"] = rng.uniform(-80, 80)
    # shuffle the order inside each family so the nearest-image matching is exercised
    for s in range(nsets):
        perm = rng.permutation(3)
        images[3 * s:3 * s + 3] = images[3 * s:3 * s + 3][perm]
    chi = C.c_double(0.0)
    found = C.c_int(0)
    locc = (C.c_int * nsets)(*loca.tolist())
    R.ref_chi_computation(C.byref(chi), C.byref(found), locc, pos, nimages.ctypes.data_as(C.POINTER(C.c_int)),
                          M.galaxy_ptr(images), nsets)
    np.savez(os.path.join(HERE, "chi.npz"), nimages=nimages, images=images, image_pos=posa,
             locimagesfound=loca, chi=np.array([chi.value]), images_found=np.array([found.value]))
    print("golden written; delensed images per source:", loca.tolist(), "chi =", chi.value)


if __name__ == "__main__":
    main()
