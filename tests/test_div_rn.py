"""The gradient pass divides by the focal lengths and by depth with an
FMA-corrected product by the reciprocal (div_rn in rt_device.hpp) instead of
IEEE division; its results are pinned here against IEEE division on the CPU
(the same three operations in C, with the C library's fused fmaf) and, on the
GPU, by the gradient bins matching the oracle bit for bit
(tests/test_gpu_batch.py)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_div_rn_equals_ieee_division(tmp_path):
    exe = str(tmp_path / "div_rn_check")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, os.path.join(HERE, "div_rn_check.c"), "-lm"],
                   check=True)
    out = subprocess.run([exe, "20000000", "160"], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0 and out.stdout.startswith("0 mismatches")
