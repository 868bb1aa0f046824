"""The full pipeline on a seeded 2000-frame sequence at 640x480 with mixed slow
(<= 1 deg, <= 7.5 mm per frame) and fast (<= 8 deg, <= 60 mm per frame)
segments of 100 frames: keyframe fusion on, learned proposals warm-starting
every frame, the plain start as fallback when the residual RMS exceeds mean +
3 sigma of the recent frames, then batched relocalisation against up to 64
keyframes (those the keyframe search finds near the estimate, then the most
recent ones), bi-objective tracking (lambda 0.5, 61440 points). Compared with the plain tracker (same
keyframe policy, starting from the previous estimate, no relocalisation):
frames/s, the fraction of frames whose motion is recovered within 0.5 deg /
5 mm, and the drift (RMS of the 100-frame relative pose error) against ground
truth."""
import numpy as np
import pytest

from recon_amd import synth
from recon_amd.cnn import PoseNet
from recon_amd.pipeline import SequenceTracker, frame_success, relative_pose_error
from tests.test_gpu_sequence import render_sequence

pytestmark = pytest.mark.gpu


def test_learned_pipeline_2000_frames():
    W, H, n = 640, 480, 2000
    print(f"case pipeline {W}x{H} n={n}", flush=True)
    g, d, gt = render_sequence(61, n, W, H, 8.0, 60.0, slow_frac=0.125, seg_len=100)
    cal = synth.calib_block_np(W, H)
    net = PoseNet.load(max_batch=64)
    learned = SequenceTracker(W, H, cal, net=net)
    P, s = learned.run(g, d, n)
    plain = SequenceTracker(W, H, cal, net=None, relocalise=False)
    Q, q = plain.run(g, d, n)
    res = {}
    for name, X, st in (("learned", P, s), ("plain", Q, q)):
        rpe_t, rpe_r = relative_pose_error(X, gt, 100)
        ok = frame_success(X, gt)
        fin = synth.pose_error(X[-1], gt[-1])
        res[name] = (rpe_t, rpe_r, ok)
        print(f"{name}: {st['fps']:.1f} frames/s, keyframes {st['keyframes']}, relocalisations "
              f"{st['relocalisations']}, frames recovered {ok * 100:.1f}%, drift per 100 frames {rpe_t:.1f} mm "
              f"{rpe_r:.2f} deg, final error {fin[0]:.2f} deg {fin[1]:.1f} mm", flush=True)
    assert res["learned"][2] > 0.95 and res["learned"][2] > res["plain"][2]
    assert res["learned"][0] < 0.25 * res["plain"][0]
