"""The serving loop's memory admission on the CPU (lenstool-hpc_amd/csrc/lt_server_policy.h, driven by
tests/cpp/server_policy_check.cpp): a caller that fetches in submission order never waits forever on
a request that can no longer be admitted (20 000 random size sequences), the budget is never
exceeded, and the plain first-fit bypass it replaces does deadlock (e.g. on 50 %, 60 %, 30 %)."""
import os
import subprocess
import tempfile

from conftest import ROOT


def test_admission_never_starves_fifo_callers():
    src = os.path.join(ROOT, "tests", "cpp", "server_policy_check.cpp")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "spc")
        subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "lenstool-hpc_amd", "csrc"), "-o", exe,
                        src], check=True, timeout=120)
        r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    f = r.stdout.split()
    assert f[0] == "rule" and f[4] == "naive", r.stdout
    dead, total, peak, naive = int(f[1]), int(f[2]), int(f[3]), int(f[5])
    assert r.returncode == 0, r.stdout
    assert dead == 0 and total > 10000 and peak <= 1000
    assert naive > 0 and f[7] == "1"
