"""CPU-side checks of the C-ABI library: it loads without a GPU, exports every function the public
headers declare, reports the absence of a GPU as a status (not a crash), and the host-side pieces
(cosmology) agree with the oracle."""
import os
import re

import pytest

import ltmodel as M
from conftest import ROOT

HEADERS = [os.path.join(ROOT, "include", "lenstool_hip", h) for h in ("lt_capi.h",)]


def declared_functions():
    names = set()
    for h in HEADERS:
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[a-z_]+[\s\*]+\**\s*(lt_[a-z0-9_]+)\s*\(", src, re.M):
            names.add(m.group(1))
    return names


def test_library_loads_and_exports_all_declared(lthip_lib):
    import lthip
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lthip_lib, n), f"{n} declared in lt_capi.h but not exported"
        assert n in lthip.SIGNATURES, f"{n} has no ctypes signature"


def test_no_gpu_is_a_status(lthip_lib):
    import ctypes as C
    import lthip
    n = C.c_int(-1)
    assert lthip_lib.lt_device_count(C.byref(n)) == 0
    if n.value == 0:
        assert lthip_lib.lt_init(0) == 9          # LT_ERR_NO_DEVICE
        assert b"GPU" in lthip_lib.lt_last_error()


def test_cosmology_matches_oracle(lthip_lib, oracle):
    import lthip
    c = M.flat_lcdm(70.0, 0.3)
    for zl, zs in ((0.4, 2.0), (0.3, 5.9), (0.05, 1.1)):
        assert abs(lthip.cosmo_dlsds(zl, zs, c) - oracle.cosmo_dlsds(zl, zs, c)) < 1e-12
    c0 = M.flat_lcdm(70.0, 0.3)
    c0.omegaX = 0.0
    assert abs(lthip_lib.lt_cosmo_observer_object(2.0, c0) - oracle.cosmo_observer_object(2.0, c0)) < 1e-14
    assert lthip.cosmo_dlsds(2.0, 1.0, c) == 0.0


def test_invalid_arguments_without_gpu(lthip_lib):
    """Argument validation runs before any device call."""
    import ctypes as C
    import lthip
    lens = M.group_lens()
    grid = M.group_grid(8)
    bad = lens.subset(range(lens.n))
    bad.type[0] = 42
    bad._soa = None
    with pytest.raises(lthip.LtError) as e:
        lthip.gradient_grid(bad, grid)
    assert e.value.status == 2                     # LT_ERR_TYPE
    with pytest.raises(lthip.LtError) as e:
        lthip_lib_rc = lthip_lib.lt_gradient_grid(None, None, C.byref(grid.frame), C.byref(lens.soa()), lens.n,
                                                  grid.dx, grid.dy, 8, 8, 0, 0, 0)
        lthip.check(lthip_lib_rc)
    assert e.value.status == 1                     # LT_ERR_INVALID
