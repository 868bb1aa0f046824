"""Write the golden vectors of tests/golden/ from the reference's own pose maths.

Runs in the build container only: it loads oracle/_ref/libref.so, which
oracle/Makefile (target `ref`) compiles from the reference sources
tracker/tracker/expm.c (EXPOKIT Pade expm, expm.c:2017-2090) and
tracker/tracker/basic_math.cpp (invertRT4, rot2Quaternion, normalizeQuaternion,
matrixMult4x4, poseDistance, inverse3x3). The fixture holds inputs and outputs
only.
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def main():
    import oracle
    oracle.build(with_reference=True)
    ref = oracle.ref_lib()
    if ref is None:
        raise SystemExit("oracle/_ref/libref.so is not built (needs the reference tree)")
    f = ctypes.POINTER(ctypes.c_float)
    d = ctypes.POINTER(ctypes.c_double)
    rng = np.random.default_rng(2016)
    out = {}

    # expm: T <- T * expm(A(x)) with translation scaled by 1e3, for twists of
    # the sizes one Gauss-Newton step produces (rotations up to ~10 deg,
    # translations up to ~100 mm after scaling), plus the zero twist
    n = 64
    xs = np.zeros((n, 6), np.float32)
    xs[:, :3] = rng.normal(0, 0.05, (n, 3))
    xs[:, 3:] = rng.normal(0, 0.03, (n, 3))
    xs[0] = 0
    xs[1, :3] = 0
    xs[2, 3:] = 0
    Ts = np.zeros((n, 16), np.float32)
    outs = np.zeros((n, 16), np.float32)
    for i in range(n):
        # a random rigid start pose (rotation from an exact axis-angle, then float)
        ax = rng.normal(size=3); ax /= np.linalg.norm(ax)
        ang = rng.uniform(0, 0.5)
        K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
        R = np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K
        T = np.eye(4); T[:3, :3] = R; T[:3, 3] = rng.normal(0, 100, 3)
        Ts[i] = T.astype(np.float32).ravel()
        t = Ts[i].copy()
        x = xs[i].copy()
        ref.ref_expm(x.ctypes.data_as(f), t.ctypes.data_as(f), ctypes.c_float(1e3))
        outs[i] = t
    out["expm_x"], out["expm_T"], out["expm_out"] = xs, Ts, outs

    # invertRT4 (double and float)
    M = np.zeros((16, 16))
    Mi = np.zeros((16, 16))
    Mf = np.zeros((16, 16), np.float32)
    Mif = np.zeros((16, 16), np.float32)
    for i in range(16):
        q = rng.normal(size=4); q /= np.linalg.norm(q)
        w, x, y, z = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                      [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                      [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
        T = np.eye(4); T[:3, :3] = R; T[:3, 3] = rng.normal(0, 500, 3)
        M[i] = T.ravel()
        a = M[i].copy(); b = np.zeros(16)
        ref.ref_invert_rt4_d(a.ctypes.data_as(d), b.ctypes.data_as(d))
        Mi[i] = b
        Mf[i] = M[i].astype(np.float32)
        a = Mf[i].copy(); b = np.zeros(16, np.float32)
        ref.ref_invert_rt4_f(a.ctypes.data_as(f), b.ctypes.data_as(f))
        Mif[i] = b
    out["inv_M"], out["inv_Mi"], out["inv_Mf"], out["inv_Mif"] = M, Mi, Mf, Mif

    # matrixMult4x4 (double)
    A = rng.normal(size=(8, 16)); B = rng.normal(size=(8, 16)); C = np.zeros((8, 16))
    for i in range(8):
        a, b, c = A[i].copy(), B[i].copy(), np.zeros(16)
        ref.ref_mat_mult4_d(a.ctypes.data_as(d), b.ctypes.data_as(d), c.ctypes.data_as(d))
        C[i] = c
    out["mul_A"], out["mul_B"], out["mul_C"] = A, B, C

    # rot2Quaternion + normalizeQuaternion (float), all four branches
    Rs = np.zeros((32, 16), np.float32)
    Q = np.zeros((32, 4), np.float32)
    QN = np.zeros((32, 4), np.float32)
    for i in range(32):
        ax = rng.normal(size=3); ax /= np.linalg.norm(ax)
        ang = [0.01, 0.2, 1.5, 2.9, 3.1][i % 5] if i < 25 else rng.uniform(0, np.pi)
        if i >= 28:
            ax = np.eye(3)[i - 28] if i < 31 else ax
            ang = 3.0
        K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
        R = np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K
        T = np.eye(4); T[:3, :3] = R
        Rs[i] = T.astype(np.float32).ravel()
        q = np.zeros(4, np.float32)
        m = Rs[i].copy()
        ref.ref_rot2quat_f(m.ctypes.data_as(f), q.ctypes.data_as(f))
        Q[i] = q
        ref.ref_normalize_quat_f(q.ctypes.data_as(f))
        QN[i] = q
    out["quat_R"], out["quat_q"], out["quat_qn"] = Rs, Q, QN

    # inverse3x3 (float): intrinsics blocks as the calibration file gives them
    # at several widths, generic matrices, and nearly singular ones (identity
    # branch)
    rng2 = np.random.default_rng(2017)
    K = np.zeros((24, 9), np.float32)
    Ki = np.zeros((24, 9), np.float32)
    for i in range(24):
        if i < 12:
            w = [160, 320, 640, 1280][i % 4]
            fx, fy = rng2.uniform(0.7, 1.0), rng2.uniform(0.7, 1.0)
            cx, cy = rng2.uniform(0.45, 0.55), rng2.uniform(0.35, 0.4)
            k = np.array([-fx * w, 0, cx * w, 0, fy * w, cy * w, 0, 0, 1])
        elif i < 20:
            k = rng2.normal(size=9)
        else:
            k = rng2.normal(size=9) * 1e-2
        K[i] = k.astype(np.float32)
        a = K[i].copy(); b = np.zeros(9, np.float32)
        ref.ref_inverse3x3_f(a.ctypes.data_as(f), b.ctypes.data_as(f))
        Ki[i] = b
    out["inv3_K"], out["inv3_Ki"] = K, Ki

    # poseDistance: translation length and the largest absolute Euler angle of
    # relative transforms (basic_math.cpp:1059-1094)
    P = np.zeros((32, 16), np.float32)
    PD = np.zeros((32, 2))
    for i in range(32):
        ax = rng2.normal(size=3); ax /= np.linalg.norm(ax)
        ang = rng2.uniform(0, [0.05, 0.5, 2.0, 3.1][i % 4])
        Kx = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
        R = np.eye(3) + np.sin(ang) * Kx + (1 - np.cos(ang)) * Kx @ Kx
        T = np.eye(4); T[:3, :3] = R; T[:3, 3] = rng2.normal(0, [1.0, 50.0, 500.0, 2000.0][i % 4], 3)
        if i == 0:
            T = np.eye(4)
        P[i] = T.astype(np.float32).ravel()
        a = P[i].copy(); dd = ctypes.c_double(); aa = ctypes.c_double()
        ref.ref_pose_distance(a.ctypes.data_as(f), ctypes.byref(dd), ctypes.byref(aa))
        PD[i] = dd.value, aa.value
    out["pd_T"], out["pd_dist_angle"] = P, PD

    np.savez(os.path.join(HERE, "ref_pose_math.npz"), **out)
    print("wrote", os.path.join(HERE, "ref_pose_math.npz"), {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
