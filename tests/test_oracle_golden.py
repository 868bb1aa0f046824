"""The oracle's pose maths against the reference's own code.

The fixture tests/golden/ref_pose_math.npz holds inputs and outputs of the
reference's tracker/tracker/expm.c (EXPOKIT Pade(6) expm with the
normalizeMat4RealSafe re-orthonormalisation, expm.c:2017-2090) and
basic_math.cpp (invertRT4, matrixMult4x4, rot2Quaternion, normalizeQuaternion),
written by tests/golden/make_golden.py from those sources compiled in place.
Where oracle/_ref/libref.so is present the same checks also run live against it.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle
from oracle import call

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_pose_math.npz"))


def test_expm_update_matches_reference_pade():
    """Closed-form SE(3) exponential + the reference's normalisation and
    composition reproduce the reference's Pade expm (FP32 outputs)."""
    for x, T0, Tref in zip(G["expm_x"], G["expm_T"], G["expm_out"]):
        T = T0.copy()
        call("or_expm_update_f", x.astype(np.float64), T, 1e3)
        # FP64 closed form and FP64 Pade differ by ~1e-16; after the float store they agree exactly
        assert np.array_equal(T, Tref), (x, T - Tref)


def test_expm_zero_twist_is_identity_increment():
    T0 = G["expm_T"][0].copy()
    assert np.all(G["expm_x"][0] == 0)
    T = T0.copy()
    call("or_expm_update_f", np.zeros(6), T, 1e3)
    assert np.array_equal(T, G["expm_out"][0])


def test_invert_rt4_bitexact():
    for M, Mi in zip(G["inv_M"], G["inv_Mi"]):
        out = np.zeros(16)
        call("or_invert_rt4_d", M.copy(), out)
        assert np.array_equal(out, Mi)
    for M, Mi in zip(G["inv_Mf"], G["inv_Mif"]):
        out = np.zeros(16, np.float32)
        call("or_invert_rt4_f", M.copy(), out)
        assert np.array_equal(out, Mi)


def test_mat4_mul_bitexact():
    for A, B, C in zip(G["mul_A"], G["mul_B"], G["mul_C"]):
        out = np.zeros(16)
        call("or_mat4_mul", A.copy(), B.copy(), out)
        assert np.array_equal(out, C)


def test_quaternion_bitexact():
    for R, q, qn in zip(G["quat_R"], G["quat_q"], G["quat_qn"]):
        a = np.zeros(4, np.float32); b = np.zeros(4, np.float32)
        call("or_rot2quat", R.copy(), a, b)
        assert np.array_equal(a, q)
        assert np.array_equal(b, qn)


@pytest.mark.skipif(oracle.ref_lib() is None, reason="oracle/_ref is built only where the reference tree exists")
def test_live_reference_expm_random():
    ref = oracle.ref_lib()
    f = ctypes.POINTER(ctypes.c_float)
    rng = np.random.default_rng(7)
    for _ in range(200):
        x = (rng.normal(0, 0.1, 6)).astype(np.float32)
        T = np.eye(4, dtype=np.float32).ravel().copy()
        T[3], T[7], T[11] = rng.normal(0, 50, 3)
        a = T.copy(); b = T.copy()
        ref.ref_expm(x.ctypes.data_as(f), a.ctypes.data_as(f), ctypes.c_float(1e3))
        call("or_expm_update_f", x.astype(np.float64), b, 1e3)
        assert np.allclose(a, b, rtol=0, atol=4e-7 * max(1.0, np.abs(a).max()))
