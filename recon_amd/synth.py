"""Seeded synthetic RGB-D inputs (see csrc/synth.cpp) as numpy arrays.

Camera poses T_wc map camera points to world points in the tracker's
convention (x right, y up, looking down -z, millimetres).
"""
import ctypes

import numpy as np

from . import synth_lib

_dp = ctypes.POINTER(ctypes.c_double)


def pinhole(width, height):
    """Intrinsics (fx, fy, cx, cy) of the workloads: fx = fy = 525 at 640 wide,
    scaled with the width (262.5 at 320), principal point at the centre."""
    f = 525.0 * width / 640.0
    return f, f, (width - 1) / 2.0, (height - 1) / 2.0


def calib_block(width, height, fx=None, fy=None, cx=None, cy=None, zmin=300.0, zmax=7000.0):
    """The 96-float calibration block for a pinhole camera (rt_calib_pinhole)."""
    from . import lib
    if fx is None:
        fx, fy, cx, cy = pinhole(width, height)
    out = (ctypes.c_float * 96)()
    rc = lib().rt_calib_pinhole(fx, fy, cx, cy, zmin, zmax, out)
    if rc != 0:
        raise ValueError("bad intrinsics")
    return np.frombuffer(out, dtype=np.float32).copy()


def calib_block_np(width, height, fx=None, fy=None, cx=None, cy=None, zmin=300.0, zmax=7000.0):
    """Same block built in numpy (no native library needed)."""
    if fx is None:
        fx, fy, cx, cy = pinhole(width, height)
    c = np.zeros(96, np.float32)
    K = np.array([-fx, 0, cx, 0, fy, cy, 0, 0, 1], np.float32)
    iK = np.array([np.float32(1) / K[0], 0, -K[2] / K[0], 0, np.float32(1) / K[4], -K[5] / K[4], 0, 0, 1], np.float32)
    c[0:9] = K; c[9:18] = iK; c[43:52] = K; c[52:61] = iK
    c[26:42] = np.eye(4, dtype=np.float32).ravel(); c[64:80] = np.eye(4, dtype=np.float32).ravel()
    c[62] = zmin; c[63] = zmax
    return c


def calib_block_rig(width, height, zmin=300.0, zmax=7000.0):
    """A calibration block of a two-camera rig: depth intrinsics KL and colour
    intrinsics KR differ, the colour camera sits 25 mm beside the depth camera
    with a small rotation (TLR, TRL = its inverse) and has radial distortion
    (k1, k2, k5). Every transform of the tracking path is then non-trivial;
    the synthetic frames themselves are rendered in one camera, so this block
    serves kernel-against-oracle comparisons, not convergence tests."""
    s = width / 640.0
    c = np.zeros(96, np.float32)

    def K_and_inv(fx, fy, cx, cy):
        K = np.array([-fx, 0, cx, 0, fy, cy, 0, 0, 1], np.float32)
        iK = np.array([np.float32(1) / K[0], 0, -K[2] / K[0], 0, np.float32(1) / K[4], -K[5] / K[4], 0, 0, 1],
                      np.float32)
        return K, iK
    KR, iKR = K_and_inv(525.0 * s, 523.0 * s, 319.5 * s + 1.7, 239.5 * s - 2.1)
    KL, iKL = K_and_inv(580.0 * s, 578.0 * s, 322.1 * s, 236.4 * s)
    c[0:9] = KR; c[9:18] = iKR; c[43:52] = KL; c[52:61] = iKL
    c[18:26] = np.array([0.04, -0.08, 0.0, 0.0, 0.01, 0, 0, 0], np.float32)
    ay, ax = np.radians(0.6), np.radians(0.3)
    Ry = np.array([[np.cos(ay), 0, np.sin(ay)], [0, 1, 0], [-np.sin(ay), 0, np.cos(ay)]])
    Rx = np.array([[1, 0, 0], [0, np.cos(ax), -np.sin(ax)], [0, np.sin(ax), np.cos(ax)]])
    T = np.eye(4)
    T[:3, :3] = Ry @ Rx
    T[:3, 3] = [25.0, 0.4, -1.2]
    c[26:42] = T.astype(np.float32).ravel()
    c[64:80] = np.linalg.inv(T).astype(np.float32).ravel()
    c[62] = zmin; c[63] = zmax
    return c


class Scene:
    def __init__(self, seed):
        self._lib = synth_lib()
        self._h = self._lib.rs_scene_create(ctypes.c_uint64(seed))

    def __del__(self):
        try:
            if self._h:
                self._lib.rs_scene_destroy(self._h)
        except Exception:
            pass

    def render(self, T_wc, width, height, noise_seed=0, noise=0.01, intrinsics=None):
        fx, fy, cx, cy = intrinsics if intrinsics else pinhole(width, height)
        T = np.ascontiguousarray(T_wc, dtype=np.float64).ravel()
        gray = np.empty((height, width), np.uint8)
        depth = np.empty((height, width), np.uint16)
        self._lib.rs_render(self._h, T.ctypes.data_as(_dp), width, height, fx, fy, cx, cy,
                            ctypes.c_uint64(noise_seed), noise, gray.ctypes.data, depth.ctypes.data)
        return gray, depth


def sample_motion(seed, theta_max_deg, t_max_mm):
    M = np.zeros(16, np.float64)
    synth_lib().rs_sample_motion(ctypes.c_uint64(seed), theta_max_deg, t_max_mm, M.ctypes.data_as(_dp))
    return M.reshape(4, 4)


def trajectory(seed, n, rot_max_deg, trans_max_mm, seg_len=0, slow_frac=1.0):
    P = np.zeros((n, 16), np.float64)
    synth_lib().rs_trajectory(ctypes.c_uint64(seed), n, rot_max_deg, trans_max_mm, seg_len, slow_frac,
                              P.ctypes.data_as(_dp))
    return P.reshape(n, 4, 4)


def make_pair(seed, width, height, theta_max_deg, t_max_mm, noise=0.01):
    """One frame pair: (ref gray, ref depth, cur gray, cur depth, Trel_gt) where
    Trel_gt maps reference-camera points into the current camera."""
    scene = Scene(seed)
    M = sample_motion(seed * 7 + 1, theta_max_deg, t_max_mm)
    T0 = np.eye(4)
    T1 = T0 @ M
    g0, d0 = scene.render(T0, width, height, noise_seed=seed * 2 + 1, noise=noise)
    g1, d1 = scene.render(T1, width, height, noise_seed=seed * 2 + 2, noise=noise)
    return g0, d0, g1, d1, np.linalg.inv(M)


def make_sequence(seed, n, width, height, rot_max_deg, trans_max_mm, noise=0.01, seg_len=0, slow_frac=1.0):
    """n frames along a smooth trajectory: gray (n,H,W) u8, depth (n,H,W) u16,
    poses (n,4,4) T_wc with poses[0] = identity."""
    scene = Scene(seed)
    poses = trajectory(seed * 13 + 5, n, rot_max_deg, trans_max_mm, seg_len, slow_frac)
    gray = np.empty((n, height, width), np.uint8)
    depth = np.empty((n, height, width), np.uint16)
    for k in range(n):
        gray[k], depth[k] = scene.render(poses[k], width, height, noise_seed=seed * 1000 + k, noise=noise)
    return gray, depth, poses


def pose_error(T_est, T_gt):
    """(rotation error in degrees, translation error in mm) of two 4x4 poses."""
    D = np.linalg.inv(np.asarray(T_gt, np.float64)) @ np.asarray(T_est, np.float64)
    R = D[:3, :3]
    # angle from both the skew part (sin) and the trace (cos): well conditioned near zero
    s = 0.5 * np.linalg.norm([R[2, 1] - R[1, 2], R[0, 2] - R[2, 0], R[1, 0] - R[0, 1]])
    c = (np.trace(R) - 1.0) / 2.0
    return float(np.degrees(np.arctan2(s, c))), float(np.linalg.norm(D[:3, 3]))


MIXED_SIZES = [(160, 120), (320, 240), (480, 360), (640, 480)]


def render_mixed_batch(n=256, seed=3, theta=3.0, tmax=50.0):
    """The mixed-size batch workload: n frame pairs with sizes drawn uniformly
    from MIXED_SIZES (numpy default_rng(seed)), one scene per pair, motions <=
    theta deg / tmax mm, rendered on the GPU into one gray and one depth
    buffer; the device is synchronised before return (the frames are rendered
    on the null stream, which a tracker's non-blocking stream does not wait
    for). Returns (pairs, gray buffer, depth buffer): per pair the rt_pair
    fields, its true Trel ("gt") and its two frames' pixel offsets ("off")."""
    from .hip import DeviceBuffer, hip
    S = synth_lib()
    sizes = [MIXED_SIZES[k] for k in np.random.default_rng(seed).integers(0, len(MIXED_SIZES), n)]
    npx = [W * H for W, H in sizes]
    off = np.concatenate([[0], np.cumsum([2 * p for p in npx])])
    g, d = DeviceBuffer(int(off[-1])), DeviceBuffer(int(off[-1]) * 2)
    pairs = []
    for i, (W, H) in enumerate(sizes):
        sc = S.rs_scene_create(ctypes.c_uint64(seed * 10000 + i))
        M = sample_motion(seed * 10000 + 7000 + i, theta, tmax)
        fx, fy, cx, cy = pinhole(W, H)
        for f, T in ((0, np.eye(4)), (1, M)):
            Tf = np.ascontiguousarray(T, np.float64).ravel()
            o = int(off[i]) + f * npx[i]
            rc = S.rs_render_device(sc, Tf.ctypes.data_as(_dp), W, H, fx, fy, cx, cy,
                                    ctypes.c_uint64(seed * 100000 + 2 * i + f), 0.01, g.ptr + o, d.ptr + 2 * o, None)
            if rc != 0:
                raise RuntimeError(f"render failed: {rc}")
        S.rs_scene_destroy(sc)
        o0, o1 = int(off[i]), int(off[i]) + npx[i]
        pairs.append(dict(ref_gray=g.ptr + o0, ref_depth=d.ptr + 2 * o0, cur_gray=g.ptr + o1, cur_depth=d.ptr + 2 * o1,
                          width=W, height=H, calib=calib_block_np(W, H), gt=np.linalg.inv(M), off=(o0, o1)))
    rc = hip().hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed: {rc}")
    return pairs, g, d
