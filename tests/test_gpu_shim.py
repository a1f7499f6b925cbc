"""The reference-signature C++ entry points (include/lenstool_hip/lenstool_gpu.hpp), driven by a
small C++ programs that use them the way a reference caller does (tests/cpp/shim_check.cpp,
tests/cpp/shim_chi_check.cpp)."""
import os
import struct
import subprocess
import tempfile

import numpy as np
import pytest

import ltmodel as M
from conftest import ROOT
from parity import EXT_SHARE

pytestmark = pytest.mark.gpu
BIN = os.path.join(ROOT, "tests", "cpp", "shim_check")


def test_reference_signatures(oracle):
    import lthip
    lthip.check(lthip.lib().lt_init(0))
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/shim_check is not built (make -C lenstool-hpc_amd/csrc)")
    lens = M.group_lens(seed=9)
    cosmo = M.flat_lcdm()
    zs = 2.0
    lens.dlsds[:] = oracle.cosmo_dlsds(0.4, 1.7, cosmo)
    lens._soa = None
    n = 72
    grid = M.group_grid(n)
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<3i5d", lens.n, n, n, -100.0, 100.0, -100.0, 100.0, zs))
            f.write(lens.type.astype("<i4").tobytes())
            for a in (lens.x, lens.y, lens.b0, lens.angle, lens.epot, lens.rcore, lens.rcut, lens.z, lens.dlsds):
                f.write(a.astype("<f8").tobytes())
        print(f"[case] shim_check group lens {lens.n} halos {n}x{n}", flush=True)
        subprocess.run([BIN, fin, fout], check=True, timeout=120)
        raw = open(fout, "rb").read()
    status = struct.unpack("<i", raw[-4:])[0]
    v = np.frombuffer(raw[:-4], dtype="<f8").reshape(-1, n, n)
    assert status == 0
    assert v.shape[0] == 29
    gx, gy = lthip.gradient_grid(lens, grid)
    np.testing.assert_array_equal(v[0], gx)
    np.testing.assert_array_equal(v[1], gy)
    H = lthip.hessian_grid(lens, grid, 1, 2)
    for k in range(4):
        np.testing.assert_array_equal(v[2 + k], H[k])
    mu = lthip.map_grid(lthip.MAP_AMPLI, 1, lens, grid, cosmo, 0.4, zs)
    np.testing.assert_array_equal(v[6], mu)
    # the foreign map function runs the epilogue on the stored planes: equal up to the rounding of
    # a = ((a + c) + (a - c)) / 2, amplified by |mu| near the critical curves
    assert (np.abs(v[7] - mu) <= 1e-12 * np.abs(mu) * np.maximum(1.0, np.abs(mu))).all()
    np.testing.assert_array_equal(v[8], lthip.map_grid(lthip.MAP_SHEAR, 3, lens, grid, cosmo, 0.4, zs))
    np.testing.assert_array_equal(v[9], lthip.map_grid(lthip.MAP_MASS, 2, lens, grid, cosmo, 0.0, zs))
    np.testing.assert_array_equal(v[10], lthip.map_grid(lthip.MAP_POT, 1, lens, grid, cosmo, 0.4, zs))
    dx, dy = lthip.map_grid(lthip.MAP_DPL, 1, lens, grid, cosmo, 0.4, zs)
    np.testing.assert_array_equal(v[11], dx)
    np.testing.assert_array_equal(v[12], dy)
    # CPU-named entry points served by the GPU: the same numbers as the GPU-named ones
    np.testing.assert_array_equal(v[13], H[0])
    np.testing.assert_array_equal(v[14], H[2])
    H0 = lthip.hessian_grid(lens, grid)
    np.testing.assert_array_equal(v[15], 0.5 * (H0[0] + H0[2]))  # amplif_grid_CPU mode 5: kappa, no dlsds
    np.testing.assert_array_equal(v[16], gx)
    pts = v[17].ravel()[:16].reshape(4, 4)
    for k in range(4):
        np.testing.assert_array_equal(pts[k], [H0[0][k, k], H0[1][k, k], H0[2][k, k], H0[3][k, k]])
    # the launchers on device data give the same planes as the host-array entry points
    g12 = lthip.gradient_grid(lens, grid, 1, 2)
    np.testing.assert_array_equal(v[18], g12[0])
    np.testing.assert_array_equal(v[19], g12[1])
    for k in range(4):
        np.testing.assert_array_equal(v[20 + k], H[k])
    K = lthip.hessian_grid(lens, grid, kmap=True)
    np.testing.assert_array_equal(v[24], K[0])
    np.testing.assert_array_equal(v[25], K[2])
    np.testing.assert_array_equal(v[26], gx)
    np.testing.assert_array_equal(v[27], H0[0])
    np.testing.assert_array_equal(v[28], dx)  # dpl mode 1: x * D_ls/D_s, as the DPL map's x plane
    # and the shim's gradient is the oracle's within the FP64 criterion
    ex, ey = oracle.gradient_grid(lens, grid, extended=True)
    _, _, ax, ay = oracle.gradient_grid(lens, grid, with_abs=True)
    cg, _ = oracle.cond_planes(lens, grid)
    assert (np.abs(v[0] - ex) <= 1e-12 * ax + EXT_SHARE * cg).all()


CHI_BIN = os.path.join(ROOT, "tests", "cpp", "shim_chi_check")


def test_reference_signatures_potential_delensing_chi(oracle):
    """potential_grid_GPU (both overloads), gradient_grid_GPU_UM, the one-point gradient, delensing
    (GPU- and CPU-named), chi_computation and the two chi^2 orchestrators, called with the reference's
    structures from C++: equal to the C-ABI calls on the same inputs (bit for bit), and the chi^2 to
    the oracle's within the gradient's tolerance."""
    import lthip
    from test_gpu_delense import _families
    lthip.check(lthip.lib().lt_init(0))
    if not os.path.exists(CHI_BIN):
        pytest.fail("tests/cpp/shim_chi_check is not built (make -C lenstool-hpc_amd/csrc)")
    lens = M.group_lens(seed=5)
    n = 64
    frame = M.grid_param(-100.0, 100.0, -100.0, 100.0)
    grid = M.Grid(frame, n, n)
    nimages, images = _families(lens, oracle, n, frame, 8, seed=23)
    nsets = len(nimages)
    src = oracle.source_barycenters(nimages, images, lens)
    gx, gy = lthip.gradient_grid(lens, grid)
    rng = np.random.default_rng(31)
    px = rng.uniform(-100, 100, 16)
    py = rng.uniform(-100, 100, 16)
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<4i4d", lens.n, n, nsets, len(images), -100.0, 100.0, -100.0, 100.0))
            f.write(lens.type.astype("<i4").tobytes())
            for a in (lens.x, lens.y, lens.b0, lens.angle, lens.epot, lens.rcore, lens.rcut, lens.z, lens.dlsds):
                f.write(a.astype("<f8").tobytes())
            f.write(nimages.astype("<i4").tobytes())
            f.write(images.tobytes())
            f.write(np.ascontiguousarray(src).tobytes())
            for a in (gx, gy, px, py):
                f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
        print(f"[case] shim_chi_check group lens {lens.n} halos {n}x{n} {nsets} families", flush=True)
        subprocess.run([CHI_BIN, fin, fout], check=True, timeout=120)
        raw = open(fout, "rb").read()
    status = struct.unpack("<i", raw[-4:])[0]
    v = np.frombuffer(raw[:-4], dtype="<f8")
    assert status == 0
    npx = n * n
    nd = nsets * M.MAXIMPERSOURCE * 2 + nsets + 1
    assert v.size == 4 * npx + 32 + 2 * nd + 6
    planes, v = v[:4 * npx].reshape(4, n, n), v[4 * npx:]
    np.testing.assert_array_equal(planes[0], lthip.potential_grid(lens, grid))
    np.testing.assert_array_equal(planes[1], lthip.potential_grid(lens, grid, 1, 2))
    g12 = lthip.gradient_grid(lens, grid, 1, 2)
    np.testing.assert_array_equal(planes[2], g12[0])
    np.testing.assert_array_equal(planes[3], g12[1])
    qx, qy = lthip.point_gradients(lens, px, py)
    np.testing.assert_array_equal(v[:16], qx)
    np.testing.assert_array_equal(v[16:32], qy)
    v = v[32:]
    st, pos, loc, numimg = lthip.delense_barycenter(nsets, n, frame, src, gx, gy)
    assert st == 0
    for k in range(2):  # delense_barycenter_GPU, then delense_barycenter
        blk = v[k * nd:(k + 1) * nd]
        np.testing.assert_array_equal(blk[:nsets * M.MAXIMPERSOURCE * 2].reshape(pos.shape), pos)
        np.testing.assert_array_equal(blk[nsets * M.MAXIMPERSOURCE * 2:-1], loc)
        assert blk[-1] == numimg
    chi, found, chi_gpu, err_gpu, chi_cpu, err_cpu = v[2 * nd:]
    assert (chi, found) == lthip.chi_computation(loc, pos, nimages, images)
    st1, chi1 = lthip.chi_barycentersource(lens, n, frame, nimages, images)
    assert st1 == 0 and err_gpu == 0 and err_cpu == 0
    assert chi_gpu == chi1 and chi_cpu == chi1
    ost, ochi = oracle.chi_barycentersource(lens, n, frame, nimages, images)
    assert ost == 0 and abs(chi1 - ochi) <= 1e-9 * abs(ochi)
