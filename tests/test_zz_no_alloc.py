"""lt_critic_device_noalloc (the contour with its row counts and offsets in a fixed scratch array
instead of a device allocation) against lt_critic_device on the same device planes: segments, kinds,
count and status identical, on two calls in a row, the second on a smaller grid (the array keeps
the first call's contents, which must not leak into the second), then once more on the first grid;
grids taller than the array allows are refused before anything runs."""
import ctypes as C

import numpy as np
import pytest

import ltmodel as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    import lthip
    lthip.check(lthip.lib().lt_init(0))
    return lthip


def _planes(hip, lens, grid, r):
    m = hip.Model(lens, hip.MAP_HESSIAN)
    kap = hip.DeviceArray(grid.nx * grid.ny)
    det = hip.DeviceArray(grid.nx * grid.ny)
    m.map(hip.MAP_FIELDS, 0, [kap.ptr, 0, 0, det.ptr, 0], grid, r=r)
    hip.sync()
    m.close()
    return kap, det


def _compare(hip, lens, grid, r, room):
    kap, det = _planes(hip, lens, grid, r)
    try:
        want = hip.critic_device(det.ptr, kap.ptr, grid, room)
        got = hip.critic_device(det.ptr, kap.ptr, grid, room, noalloc=True)
        assert got[3] == want[3] and got[2] == want[2]
        np.testing.assert_array_equal(got[0], want[0])
        np.testing.assert_array_equal(got[1], want[1])
        return want[2]
    finally:
        kap.close()
        det.close()


def test_noalloc_contour_equals_allocating_one(hip):
    lens = M.cluster_lens(seed=5, ngal=60)
    big = M.Grid(M.grid_param(-150.0, 150.0, -150.0, 150.0), 700, 600)
    small = M.Grid(M.grid_param(-60.0, 40.0, -50.0, 45.0), 157, 131)
    print(f"[case] no-alloc contour: {lens.n} halos, grids 700x600 then 157x131 then 700x600", flush=True)
    n_big = _compare(hip, lens, big, 0.7, 1 << 16)
    n_small = _compare(hip, lens, small, 0.7, 1 << 16)
    assert n_big > 100 and n_small > 10
    # overflow: the first 10 segments and LT_ERR_OVERFLOW, as the allocating entry point
    assert _compare(hip, lens, big, 0.7, 10) == n_big


def test_noalloc_refuses_tall_grids(hip):
    lib = hip.lib()
    rows = lib.lt_critic_device_noalloc  # bound: exists and is exported
    ny = 16384 + 2
    g = hip.LtGridDesc(0.0, 0.0, 1.0, 1.0, 4, ny, 0, 0, 4)
    total = C.c_longlong(-1)
    # device pointers are never touched: the size check comes first
    st = rows(C.c_void_p(1 << 20), None, C.byref(g), None, None, 0, C.byref(total), None)
    assert st == 6 and total.value == -1  # LT_ERR_CAPACITY
    assert b"no-alloc" in lib.lt_last_error()
