"""The delensing scan's cull (lenstool-hpc_amd/csrc/lt_delense_geom.h) on the host, compiled from the
same header as the device kernel: whenever it skips a (cell, source) pair, the reference's triangle
tests find no hit -- over lens-like cells, collinear and collapsed mapped triangles, huge gradients,
sources on corners and edges, NaN and infinity -- and the two floating-point error bounds its proof
uses hold against __float128 arithmetic. Host only; the device kernel's results are compared with the
oracle bit for bit by the delensing GPU tests."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC) and shutil.which("hipcc") is None, reason="hipcc not available")
def test_cull_is_sound(tmp_path):
    hipcc = HIPCC if os.path.exists(HIPCC) else shutil.which("hipcc")
    exe = tmp_path / "delense_cull_check"
    subprocess.run([hipcc, "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "lenstool-hpc_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "delense_cull_check.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    r = subprocess.run([str(exe), "1000000"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "unsound 0" in r.stdout and r.stdout.strip().endswith("OK")
