"""Keyframe tracking of a frame sequence with learned warm starts and batched
relocalisation, composed from the library's C ABI.

Per frame: the pose network proposes the current camera relative to the
keyframe (from the raw keyframe and current frames) and the single-pair
tracker (bi-objective by default) runs Gauss-Newton from that proposal. The
plain method stays the fallback and the arbiter: when the frame's weighted residual RMS
exceeds mean + reloc_sigma * std of the recent frames, the frame is also
tracked from the plain start (the previous estimate) and the lower RMS wins.
If the RMS still spikes, the frame is tracked against up to reloc_keyframes
keyframes in one batched launch from learned proposals: first those the
reference's keyframe search finds near the current estimate (pose distance
within reloc_search = (mm, degrees), ranked by the image displacement of a
test grid, KeyFrameModel::findSimilarKeyFrames), then the most recent ones; the best few are re-tracked by a second single-pair tracker with the
same settings (so that their RMS compares with the frame's), and tracking moves
to the best of them when it fits better and agrees with the current estimate
within reloc_gate (degrees, mm). The first frames after a keyframe are fused
into it and the keyframe is then refreshed; a new keyframe (the current frame)
is taken when the pose relative to the keyframe exceeds the distance / angle
thresholds or after max_interval frames.

With net=None and relocalise=False the same loop is the plain tracker: each
frame starts from the previous estimate (from identity after a keyframe switch),
as PhotoTracker does.
"""
import collections
import time

import numpy as np

from . import BIOBJECTIVE, PHOTOMETRIC
from .cnn import pair_array
from .hip import DeviceBuffer, hip
from .tracker import BatchTracker, PhotoTracker
from .trajectory import grid_object, similar_keyframes


def _angle_dist(T):
    c = np.clip((np.trace(T[:3, :3]) - 1.0) / 2.0, -1.0, 1.0)
    return float(np.degrees(np.arccos(c))), float(np.linalg.norm(T[:3, 3]))


def _exceeded(T, dist_mm, angle_deg):
    a, t = _angle_dist(T)
    return t > dist_mm or a > angle_deg


class SequenceTracker:
    def __init__(self, width, height, calib, net=None, mode=BIOBJECTIVE, lam=0.5, n_select=61440, fuse_frames=5,
                 max_interval=10, dist_mm=60.0, angle_deg=5.0, relocalise=True, reloc_keyframes=64, reloc_sigma=3.0,
                 reloc_window=30, reloc_points=8192, reloc_verify=3, reloc_gate=(20.0, 300.0),
                 reloc_search=(1000.0, 45.0), reloc_history=1024):
        self.W, self.H = width, height
        self.calib = calib
        self.net = net
        self.mode, self.lam = mode, lam
        self.fuse_frames, self.max_interval = fuse_frames, max_interval
        self.dist_mm, self.angle_deg = dist_mm, angle_deg
        self.relocalise = relocalise and net is not None
        self.reloc_keyframes, self.reloc_sigma, self.reloc_window = reloc_keyframes, reloc_sigma, reloc_window
        self.reloc_verify, self.reloc_gate = reloc_verify, reloc_gate
        self.reloc_search, self.reloc_history = reloc_search, reloc_history
        self.grid = grid_object()
        self.tr = PhotoTracker(width, height, calib, mode=mode, n_select=n_select, lam=lam)
        if net is not None:
            self.x = DeviceBuffer(2 * net.input_elems)
            self.o = DeviceBuffer(24)
        if self.relocalise:
            self.bt = BatchTracker(reloc_keyframes, width, height, mode=PHOTOMETRIC, n_select=reloc_points)
            self.tr2 = PhotoTracker(width, height, calib, mode=mode, n_select=n_select, lam=lam)

    def _frame(self, g, d, k):
        npx = self.W * self.H
        return g.ptr + k * npx, d.ptr + k * npx * 2

    def _propose(self, g, d, ref, cur):
        """Network proposal of the current camera pose relative to the keyframe."""
        rg, rd = self._frame(g, d, ref)
        cg, cd = self._frame(g, d, cur)
        arr = pair_array([dict(ref_gray=rg, ref_depth=rd, cur_gray=cg, cur_depth=cd, width=self.W, height=self.H)])
        self.net.make_inputs(arr, 1, self.x.ptr)
        self.net.forward(1, self.x.ptr, self.o.ptr)
        hip().hipDeviceSynchronize()
        Trel = self.net.target_to_pose(self.o.to_array(np.float32, 6))[0]
        return np.linalg.inv(Trel)

    def _set_keyframe(self, g, d, k):
        cg, cd = self._frame(g, d, k)
        self.tr.set_reference_points(cg, cd, mode=self.mode, lam=self.lam, force=True, on_device=True)

    def _track(self, g, d, k, T_init, tr=None):
        tr = tr or self.tr
        cg, cd = self._frame(g, d, k)
        tr.set_current_image(cg, cd, on_device=True)
        if T_init is None:
            tr.track()
        else:
            tr.track_from(T_init)
        return tr.get_relative_pose().astype(np.float64), tr.residual_rms()

    def _candidates(self, keyframes, G_est):
        """Keyframes to relocalise against: those near G_est by the keyframe
        search, most similar first, then the most recent others."""
        kfs = list(keyframes)
        poses = np.array([G for (_, G) in kfs], np.float32)
        near, _ = similar_keyframes(G_est.astype(np.float32), poses, self.reloc_search[0], self.reloc_search[1],
                                    self.calib[:9], self.grid)
        chosen = [int(i) for i in near[:self.reloc_keyframes]]
        for i in range(len(kfs) - 1, -1, -1):
            if len(chosen) >= self.reloc_keyframes:
                break
            if i not in chosen:
                chosen.append(i)
        return [kfs[i] for i in chosen]

    def _relocalise(self, g, d, k, keyframes, G_est):
        """Track frame k against the candidate keyframes in one batched launch, then
        re-track the best reloc_verify of them with the second single-pair
        tracker; returns (keyframe index, its global pose, pose relative to it,
        rms) of the best one that agrees with the estimate G_est, or None."""
        cand = self._candidates(keyframes, G_est)
        cg, cd = self._frame(g, d, k)
        pairs = []
        for (kf, _) in cand:
            rg, rd = self._frame(g, d, kf)
            pairs.append(dict(ref_gray=rg, ref_depth=rd, cur_gray=cg, cur_depth=cd, width=self.W, height=self.H,
                              calib=self.calib))
        self.bt.prepare_only(self.bt.prepare(pairs))
        self.bt.propose(self.net)
        Trel, T, rms, st = self.bt.retrack(None, len(pairs))
        rms = np.where(st == 0, rms, np.inf)
        best = None
        for j in np.argsort(rms)[:self.reloc_verify]:
            if not np.isfinite(rms[j]):
                break
            kf, G_j = cand[j]
            rg, rd = self._frame(g, d, kf)
            self.tr2.set_reference_points(rg, rd, mode=self.mode, lam=self.lam, force=True, on_device=True)
            T_j, r_j = self._track(g, d, k, T[j], self.tr2)
            a, t = _angle_dist(np.linalg.inv(G_est) @ (G_j @ T_j))
            if a > self.reloc_gate[0] or t > self.reloc_gate[1]:
                continue
            if best is None or r_j < best[3]:
                best = (kf, G_j, T_j, r_j)
        return best

    def run(self, g, d, n):
        """Global poses (n, 4, 4) of the n frames held in the device buffers g
        (8-bit gray) and d (16-bit depth), frame k at offset k*W*H, and
        statistics."""
        t0 = time.perf_counter()
        poses = np.zeros((n, 4, 4))
        poses[0] = np.eye(4)
        self._set_keyframe(g, d, 0)
        kf, G_kf = 0, np.eye(4)
        keyframes = collections.deque([(0, np.eye(4))], maxlen=self.reloc_history)
        hist = collections.deque(maxlen=self.reloc_window)
        n_kf, n_reloc = 1, 0
        trace = []  # per frame: (keyframe, rms, 0 first estimate / 1 plain fallback / 2 relocalised)
        T_last = np.eye(4)  # the previous frame relative to the keyframe
        for k in range(1, n):
            T_init = self._propose(g, d, kf, k) if self.net is not None else None
            T, rms = self._track(g, d, k, T_init)
            spike = len(hist) >= 10 and rms > np.mean(hist) + self.reloc_sigma * np.std(hist)
            how = 0  # 0 first estimate, 1 plain fallback, 2 relocalised
            if spike and self.net is not None:
                T2, rms2 = self._track(g, d, k, T_last)
                if rms2 < rms:
                    T, rms, how = T2, rms2, 1
                spike = rms > np.mean(hist) + self.reloc_sigma * np.std(hist)
            if spike and self.relocalise:
                best = self._relocalise(g, d, k, keyframes, G_kf @ T)
                if best is not None and best[3] < rms:
                    j, G_j, T_j, _ = best
                    self._set_keyframe(g, d, j)
                    kf, G_kf = j, G_j
                    T, rms = self._track(g, d, k, T_j)
                    n_reloc += 1
                    how = 2
            if how:
                self.tr.reset_to_transform(T)  # the device pose is the chosen estimate
            trace.append((kf, rms, how))
            hist.append(rms)
            G = G_kf @ T
            poses[k] = G
            T_last = T
            j = k - kf
            if 1 <= j <= self.fuse_frames:
                self.tr.fuse_current()
                if j == self.fuse_frames:
                    self.tr.refresh_keyframe()
            if j >= self.max_interval or _exceeded(T, self.dist_mm, self.angle_deg):
                self._set_keyframe(g, d, k)
                kf, G_kf = k, G
                keyframes.append((k, G))
                n_kf += 1
                T_last = np.eye(4)
        dt = time.perf_counter() - t0
        return poses, {"seconds": dt, "fps": n / dt, "keyframes": n_kf, "relocalisations": n_reloc, "trace": trace}


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
