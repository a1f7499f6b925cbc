"""The device FP64 atan2 and log of the gradient and potential kernels (lt_math.h atan2_lt, its
x > 0 form atan2p_lt, log_lt) against long double on the host: within 4 units in the last place over
4M random inputs (measured 2.7, 2.4); and the wavefront maximum wave_max64 equal to the host's on
every lane of 65 536 wavefronts."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "math_accuracy")


def test_atan2_log_ulp():
    assert os.path.exists(BIN), "tests/cpp/math_accuracy is not built"
    print("[case] atan2_lt / log_lt accuracy, 4M inputs", flush=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    m = re.search(r"atan2_lt max error ([0-9.]+) ulp, log_lt max error ([0-9.]+) ulp .*atan2p_lt max error ([0-9.]+) ulp",
                  r.stdout)
    assert m and float(m.group(1)) < 4 and float(m.group(2)) < 4 and float(m.group(3)) < 4
    # the wavefront maximum of the delensing scan's tile test (lt_math.h wave_max64) on every lane
    assert "wave_max64 lanes not equal to their wavefront's maximum: 0 of" in r.stdout
