"""The reference's FITS writer signatures (include/lenstool_hip/lenstool_gpu.hpp), called from a
plain g++ program: file names as the reference forms them (<dir>/<name>_<k>.fits and
<dir>/<name>.fits), relative or absolute header by ra = dec = 0, and files equal, DATE aside, to
lt_fits_write's. Host only."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT


@pytest.fixture(scope="module")
def lt():
    import lthip
    return lthip


def _strip_date(raw):
    out = bytearray(raw)
    for k in range(0, 2880 * 2, 80):
        if raw[k:k + 8] == b"DATE    ":
            out[k:k + 80] = b" " * 80
    return bytes(out)


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_fits_shims(lt, tmp_path):
    exe = tmp_path / "fits_shim_check"
    lib = os.path.join(ROOT, "lenstool-hpc_amd")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "fits_shim_check.cpp"), "-o", str(exe), "-L" + lib,
                    "-llenstool_hip", "-Wl,-rpath," + lib], check=True)
    r = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    n = 7
    img = (0.25 * np.arange(n * n) - 3.0).reshape(n, n)
    want = tmp_path / "want_rel.fits"
    lt.fits_write(want, img, -30.0, 30.0, -20.0, 40.0)
    for name in ("rel_3.fits", "img.fits"):
        assert _strip_date(open(tmp_path / name, "rb").read()) == _strip_date(open(want, "rb").read()), name
    want = tmp_path / "want_abs.fits"
    lt.fits_write(want, img, -30.0, 30.0, -20.0, 40.0, ra=3.586259, dec=-30.4)
    assert _strip_date(open(tmp_path / "abs.fits", "rb").read()) == _strip_date(open(want, "rb").read())
    # module_writeFits_ImageAbsoluteCoordinates writes the absolute header even at ra = dec = 0
    raw = open(tmp_path / "imgabs.fits", "rb").read()
    assert b"RA---TAN" in raw[:2880] and b"EQUINOX" in raw[:2880]
