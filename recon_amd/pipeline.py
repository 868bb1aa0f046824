"""Keyframe tracking of a frame sequence with learned warm starts and batched
relocalisation: the Python face of include/rt_sequence.h (the driver itself
is C++ in the library, recon_amd/csrc/rt_sequence.hip).

Per frame: the pose network proposes the current camera relative to the
keyframe and the single-pair tracker (bi-objective by default) runs
Gauss-Newton from that proposal. The plain method stays the fallback and the
arbiter: when the frame's weighted residual RMS exceeds mean + reloc_sigma *
std of the recent frames, the frame is also tracked from the plain start (the
previous estimate) and the lower RMS wins. If the RMS still spikes, the frame
is tracked against up to reloc_keyframes keyframes in one batched launch from
learned proposals: first those the reference's keyframe search finds near the
current estimate (pose distance within reloc_search = (mm, degrees), ranked by
the image displacement of a test grid, KeyFrameModel::findSimilarKeyFrames),
then the most recent ones; the best few are re-tracked by a second single-pair
tracker with the same settings (so that their RMS compares with the frame's),
and tracking moves to the best of them when it fits better and agrees with the
current estimate within reloc_gate (degrees, mm). The first frames after a
keyframe are fused into it and the keyframe is then refreshed; a new keyframe
(the current frame) is taken when the pose relative to the keyframe exceeds
the distance / angle thresholds or after max_interval frames.

With net=None and relocalise=False the same loop is the plain tracker: each
frame starts from the previous estimate (from identity after a keyframe switch),
as PhotoTracker does.
"""
import ctypes
import time

import numpy as np

from . import BIOBJECTIVE, check, lib


class RtSequenceParams(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int), ("lambda_", ctypes.c_float), ("n_select", ctypes.c_int),
                ("fuse_frames", ctypes.c_int), ("max_interval", ctypes.c_int), ("dist_mm", ctypes.c_float),
                ("angle_deg", ctypes.c_float), ("relocalise", ctypes.c_int), ("reloc_keyframes", ctypes.c_int),
                ("reloc_sigma", ctypes.c_float), ("reloc_window", ctypes.c_int), ("reloc_points", ctypes.c_int),
                ("reloc_verify", ctypes.c_int), ("reloc_gate_deg", ctypes.c_float), ("reloc_gate_mm", ctypes.c_float),
                ("reloc_search_mm", ctypes.c_float), ("reloc_search_deg", ctypes.c_float),
                ("reloc_history", ctypes.c_int)]


class SequenceTracker:
    def __init__(self, width, height, calib, net=None, mode=BIOBJECTIVE, lam=0.5, n_select=61440, fuse_frames=5,
                 max_interval=10, dist_mm=60.0, angle_deg=5.0, relocalise=True, reloc_keyframes=64, reloc_sigma=3.0,
                 reloc_window=30, reloc_points=8192, reloc_verify=3, reloc_gate=(20.0, 300.0),
                 reloc_search=(1000.0, 45.0), reloc_history=1024, device=0):
        self.L = lib()
        self.W, self.H = width, height
        self.calib = np.ascontiguousarray(calib, np.float32)
        self.net = net  # borrowed by the library: kept alive here
        p = RtSequenceParams()
        self.L.rt_sequence_default_params(ctypes.byref(p))
        p.mode, p.lambda_, p.n_select = mode, lam, n_select
        p.fuse_frames, p.max_interval, p.dist_mm, p.angle_deg = fuse_frames, max_interval, dist_mm, angle_deg
        p.relocalise, p.reloc_keyframes, p.reloc_sigma = int(relocalise), reloc_keyframes, reloc_sigma
        p.reloc_window, p.reloc_points, p.reloc_verify = reloc_window, reloc_points, reloc_verify
        p.reloc_gate_deg, p.reloc_gate_mm = reloc_gate
        p.reloc_search_mm, p.reloc_search_deg = reloc_search
        p.reloc_history = reloc_history
        self.params = p
        h = ctypes.c_void_p()
        check(self.L.rt_sequence_create(device, width, height,
                                        self.calib.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                        net.h if net is not None else None, ctypes.byref(p), ctypes.byref(h)),
              "rt_sequence_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.rt_sequence_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, g, d, n):
        """Global poses (n, 4, 4) of the n frames held in the device buffers g
        (8-bit gray) and d (16-bit depth), frame k at offset k*W*H, and
        statistics."""
        poses = np.zeros((n, 16), np.float64)
        trace = np.zeros((n, 2), np.int32)
        rms = np.zeros(n, np.float32)
        stats = np.zeros(2, np.int32)
        t0 = time.perf_counter()
        check(self.L.rt_sequence_run(self.h, g.ptr, d.ptr, n, poses.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                     trace.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                     rms.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                     stats.ctypes.data_as(ctypes.POINTER(ctypes.c_int))), "rt_sequence_run")
        dt = time.perf_counter() - t0
        tr = [(int(trace[k, 0]), float(rms[k]), int(trace[k, 1])) for k in range(1, n)]
        return poses.reshape(n, 4, 4), {"seconds": dt, "fps": n / dt, "keyframes": int(stats[0]),
                                        "relocalisations": int(stats[1]), "trace": tr}


def relative_pose_error(P, GT, delta=100):
    """RMS translation (mm) and rotation (deg) of the relative pose error over
    windows of delta frames."""
    et, er = [], []
    for k in range(len(P) - delta):
        Dg = np.linalg.inv(GT[k]) @ GT[k + delta]
        De = np.linalg.inv(P[k]) @ P[k + delta]
        E = np.linalg.inv(Dg) @ De
        et.append(np.linalg.norm(E[:3, 3]))
        c = np.clip((np.trace(E[:3, :3]) - 1.0) / 2.0, -1.0, 1.0)
        er.append(np.degrees(np.arccos(c)))
    return float(np.sqrt(np.mean(np.square(et)))), float(np.sqrt(np.mean(np.square(er))))


def frame_success(P, GT, rot_deg=0.5, trans_mm=5.0):
    """Fraction of frames whose frame-to-frame motion is estimated within the
    given bounds of the true motion."""
    ok = 0
    for k in range(1, len(P)):
        E = np.linalg.inv(np.linalg.inv(GT[k - 1]) @ GT[k]) @ (np.linalg.inv(P[k - 1]) @ P[k])
        c = np.clip((np.trace(E[:3, :3]) - 1.0) / 2.0, -1.0, 1.0)
        ok += np.degrees(np.arccos(c)) <= rot_deg and np.linalg.norm(E[:3, 3]) <= trans_mm
    return ok / (len(P) - 1)
