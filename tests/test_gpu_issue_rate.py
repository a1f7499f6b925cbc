"""VALU issue rates of the instruction forms the map kernels are built from (tests/cpp/issue_rate.hip),
measured on the GPU: the figures behind the FP32-compensated kernels' ceiling. A non-packed FP32 FMA
with all-VGPR operands issues faster than an FP64 FMA, but at most twice as fast; with one SGPR
operand (a halo constant in a scalar register) it issues slower than with VGPRs."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "issue_rate")


def test_issue_rates():
    assert os.path.exists(BIN), "tests/cpp/issue_rate is not built"
    print("[case] VALU issue rates", flush=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    rate = {m.group(1): float(m.group(2)) for m in re.finditer(r"rate (\S+)\s+([0-9.]+) wave-instr", r.stdout)}
    assert len(rate) == 8 and all(v > 0 for v in rate.values())
    fp32 = max(rate[f"fp32_fma_vgpr_{c}"] for c in (1, 2, 4, 8))
    assert rate["fp64_fma_vgpr_4"] < fp32 <= 2.2 * rate["fp64_fma_vgpr_4"]
    assert rate["fp32_fma_sgpr_4"] < rate["fp32_fma_vgpr_4"]
