"""Keyframe switching of the single-pair tracker on the GPU (referenceOutdated
/ poseDifferenceExceeded, phototracker.cpp:268-337): with reference updating on,
rt_set_reference(force = 0) replaces the keyframe exactly when the oracle's
pose-difference rule fires on the tracker's own relative pose; incremental mode
replaces it every time; with updating off only the first call does."""
import ctypes

import numpy as np
import pytest

import oracle
from recon_amd import PHOTOMETRIC, synth
from recon_amd.tracker import PhotoTracker
from tests.test_gpu_sequence import host_frames, render_sequence

pytestmark = pytest.mark.gpu


def _exceeded(T, tthr, athr):
    L = oracle.lib()
    L.or_pose_difference_exceeded.restype = ctypes.c_int
    t = np.ascontiguousarray(T, np.float32).ravel()
    return bool(L.or_pose_difference_exceeded(oracle.ptr(t), ctypes.c_float(tthr), ctypes.c_float(athr)))


def test_reference_updating_follows_the_pose_rule():
    W, H, n = 160, 120, 30
    print(f"case keyframe switching {W}x{H} n={n}", flush=True)
    g, d, gt = render_sequence(51, n, W, H, 1.5, 20.0)
    gh, dh = host_frames(g, d, n, W, H)
    cal = synth.calib_block_np(W, H)
    tthr, athr = 40.0, 3.0
    tr = PhotoTracker(W, H, cal, mode=PHOTOMETRIC, n_select=2048, reference_updating=1, pose_dist_threshold=tthr,
                      pose_angle_threshold=athr)
    assert tr.set_reference_points(gh[0], dh[0], mode=PHOTOMETRIC) == 1
    switches = 0
    for k in range(1, n):
        tr.set_current_image(gh[k], dh[k])
        assert tr.track() == 1
        T = tr.get_relative_pose()
        expect = _exceeded(T, tthr, athr)
        got = tr.set_reference_points(gh[k], dh[k], mode=PHOTOMETRIC)
        assert got == int(expect), (k, got, expect, T)
        switches += got
    assert switches >= 2
    assert tr.num_reference_updates() == switches + 1
    # the keyframe chain still follows the true trajectory (a keyframe every few
    # frames at 160x120: some drift)
    e = synth.pose_error(tr.get_global_pose(), gt[n - 1])
    assert e[0] < 1.0 and e[1] < 50.0, e


def test_incremental_and_updating_off():
    W, H, n = 160, 120, 6
    print("case keyframe incremental / off", flush=True)
    g, d, gt = render_sequence(52, n, W, H, 1.5, 20.0)
    gh, dh = host_frames(g, d, n, W, H)
    cal = synth.calib_block_np(W, H)
    off = PhotoTracker(W, H, cal, mode=PHOTOMETRIC, n_select=2048, reference_updating=0)
    inc = PhotoTracker(W, H, cal, mode=PHOTOMETRIC, n_select=2048, incremental=1)
    assert off.set_reference_points(gh[0], dh[0], mode=PHOTOMETRIC) == 1
    assert inc.set_reference_points(gh[0], dh[0], mode=PHOTOMETRIC) == 1
    for k in range(1, n):
        for tr in (off, inc):
            tr.set_current_image(gh[k], dh[k])
            assert tr.track() == 1
        assert off.set_reference_points(gh[k], dh[k], mode=PHOTOMETRIC) == 0
        assert inc.set_reference_points(gh[k], dh[k], mode=PHOTOMETRIC) == 1
    assert off.num_reference_updates() == 1 and inc.num_reference_updates() == n
    # incremental tracking chains frame-to-frame increments to the same trajectory
    for tr in (off, inc):
        e = synth.pose_error(tr.get_global_pose(), gt[n - 1])
        assert e[0] < 0.5 and e[1] < 10.0, e
