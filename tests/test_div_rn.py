"""The gradient pass divides by the focal lengths and by depth with an
FMA-corrected product by the reciprocal (div_rn in rt_device.hpp) instead of
IEEE division; its results are pinned here against IEEE division on the CPU
(the same three operations in C, with the C library's fused fmaf) and, on the
GPU, by the gradient bins matching the oracle bit for bit
(tests/test_gpu_batch.py). The reciprocal of the depth (rcp_depth: the
hardware reciprocal and one FMA Newton step) is compared with IEEE float32
division over its whole domain, every negated 16-bit depth value, on the GPU."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
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


@pytest.mark.gpu
def test_depth_reciprocal_equals_ieee_division():
    """rcp_depth(-u) == 1.0f / -u bit for bit for u = 1 .. 65535 (numpy's
    float32 division is IEEE, round to nearest)."""
    from recon_amd import lib
    print("case depth reciprocals u = 1 .. 65535", flush=True)
    n = 65535
    out = np.zeros(n, np.float32)
    rc = lib().rt_debug_depth_reciprocals(0, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n)
    assert rc == 0
    ref = np.float32(1.0) / -np.arange(1, n + 1, dtype=np.float32)
    bad = np.nonzero(out.view(np.uint32) != ref.view(np.uint32))[0]
    assert bad.size == 0, f"{bad.size} mismatches, first at u = {bad[:8] + 1}"
