"""Trajectory and schedule files and the similar-keyframe search through the C
ABI (include/rt_trajectory.h)."""
import ctypes

import numpy as np

from . import lib

_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class ScheduleEntry(ctypes.Structure):
    _fields_ = [("frame", ctypes.c_int), ("mode", ctypes.c_int), ("lambda_", ctypes.c_float), ("error", ctypes.c_float)]


def _check(rc, what):
    if rc != 0:
        msg = lib().rt_last_error()
        raise RuntimeError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def _poses(a):
    return np.ascontiguousarray(np.asarray(a, np.float32).reshape(-1, 16))


def write_trajectory(path, poses, transpose=False, inverse=False):
    """Write poses (n, 4, 4) as the reference's trajectory file (canonised to
    the first pose)."""
    p = _poses(poses)
    _check(lib().rt_trajectory_write(str(path).encode(), len(p), p.ctypes.data_as(_fp), int(transpose), int(inverse)),
           "rt_trajectory_write")


def read_trajectory(path, transpose=False, inverse=False, max_poses=100000):
    out = np.zeros((max_poses, 16), np.float32)
    n = ctypes.c_int()
    _check(lib().rt_trajectory_read(str(path).encode(), int(transpose), int(inverse), max_poses, out.ctypes.data_as(_fp),
                                    ctypes.byref(n)), "rt_trajectory_read")
    return out[:n.value].reshape(-1, 4, 4)


def read_schedule(path, max_entries=100000):
    arr = (ScheduleEntry * max_entries)()
    n = ctypes.c_int()
    _check(lib().rt_schedule_read(str(path).encode(), max_entries, ctypes.cast(arr, ctypes.c_void_p), ctypes.byref(n)),
           "rt_schedule_read")
    return arr, n.value


def schedule_lookup(entries, n, frame):
    mode, lam = ctypes.c_int(), ctypes.c_float()
    _check(lib().rt_schedule_lookup(ctypes.cast(entries, ctypes.c_void_p), n, frame, ctypes.byref(mode),
                                    ctypes.byref(lam)), "rt_schedule_lookup")
    return mode.value, lam.value


def pose_distance(dT):
    d, a = ctypes.c_double(), ctypes.c_double()
    m = np.ascontiguousarray(np.asarray(dT, np.float32).ravel())
    _check(lib().rt_pose_distance(m.ctypes.data_as(_fp), ctypes.byref(d), ctypes.byref(a)), "rt_pose_distance")
    return d.value, a.value


def grid_object(nx=30, ny=30, nz=5, size=1000.0, view_min=2000.0, z_range=3000.0, fov_x_deg=60.0):
    """The keyframe search's test object (the reference's defaults: a 30x30x5
    grid, 2 m square, 2-5 m in front of the camera)."""
    out = np.zeros(3 * nx * ny * nz, np.float32)
    _check(lib().rt_grid_object(nx, ny, nz, size, view_min, z_range, fov_x_deg, out.ctypes.data_as(_fp)),
           "rt_grid_object")
    return out.reshape(-1, 3)


def similar_keyframes(T, key_poses, max_dist, max_angle, K, grid):
    """(indices into key_poses, weights) of the keyframes near pose T, most
    similar first."""
    kp = _poses(key_poses)
    t = np.ascontiguousarray(np.asarray(T, np.float32).ravel())
    k = np.ascontiguousarray(np.asarray(K, np.float32).ravel()[:9])
    g = np.ascontiguousarray(np.asarray(grid, np.float32).ravel())
    order = np.zeros(max(1, len(kp)), np.int32)
    w = np.zeros(max(1, len(kp)), np.float32)
    n = ctypes.c_int()
    _check(lib().rt_keyframes_similar(t.ctypes.data_as(_fp), len(kp), kp.ctypes.data_as(_fp), max_dist, max_angle,
                                      k.ctypes.data_as(_fp), g.ctypes.data_as(_fp), len(g) // 3,
                                      order.ctypes.data_as(_ip), w.ctypes.data_as(_fp), ctypes.byref(n)),
           "rt_keyframes_similar")
    return order[:n.value], w[:n.value]
