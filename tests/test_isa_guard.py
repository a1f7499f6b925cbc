"""Guard against a code-generation fault seen with this toolchain: when one kernel chose between
several output pointers at run time, the compiler left the base register of the merged store
undefined on one path (`implicit-def: $sgpr..` in the ISA) and the store faulted. The kernels are
written so that no such merge exists; this test compiles them for gfx950 and checks the ISA."""
import os
import re
import shutil
import subprocess
import tempfile

import pytest

from conftest import ROOT

CSRC = os.path.join(ROOT, "lenstool-hpc_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
@pytest.mark.parametrize("src", ["lt_kernels64.hip", "lt_kernels32.hip", "lt_critic.hip"])
def test_no_undefined_scalar_registers_in_kernels(src):
    path = os.path.join(CSRC, src)
    if not os.path.exists(path):
        pytest.skip(f"{src} not present")
    d = tempfile.mkdtemp()
    try:
        out = os.path.join(d, "k.s")
        subprocess.run([HIPCC, "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                        "-O3", "-std=c++17", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
                        "--cuda-device-only", "-S", path, "-o", out], check=True, capture_output=True)
        asm = open(out).read()
    finally:
        shutil.rmtree(d)
    kernels = re.split(r"\n(?=_ZN2lt\w*:)", asm)
    checked = 0
    for k in kernels:
        m = re.match(r"(_ZN2lt\w*):", k)
        if not m or "global_store" not in k:
            continue
        checked += 1
        assert "implicit-def: $sgpr" not in k, f"{m.group(1)}: undefined scalar register in a storing kernel"
    assert checked > 0
