"""The single-pair case at 320x240 (fx = fy = 262.5): one seeded pair, rotation
<= 2 deg, translation <= 30 mm, 3 levels x 10 Gauss-Newton iterations,
256-bin MAD Tukey weights, FP32 with FP64 solve. The GPU result agrees with the
FP32 oracle within 1e-5 rad / 1e-2 mm and with the FP64 oracle within 1e-2 deg
/ 0.1 mm, and recovers the true motion; frames/s of the GPU tracker and the
time of the FP64 oracle are printed."""
import time

import numpy as np
import pytest

import oracle
from recon_amd import PHOTOMETRIC, synth
from recon_amd.tracker import PhotoTracker

pytestmark = pytest.mark.gpu


def test_single_pair_320x240():
    W, H = 320, 240
    print(f"case single pair {W}x{H}", flush=True)
    g0, d0, g1, d1, Tgt = synth.make_pair(1, W, H, 2.0, 30.0)
    cal = synth.calib_block_np(W, H)
    kw = dict(mode=PHOTOMETRIC, n_select=8192, iters=(10, 10, 10, 0), mad_bins=256)
    tr = PhotoTracker(W, H, cal, **kw)
    tr.set_reference_points(g0, d0, mode=PHOTOMETRIC, force=True)
    tr.set_current_image(g1, d1)
    assert tr.track()
    T = tr.trel().astype(np.float64)
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        tr.set_reference_points(g0, d0, mode=PHOTOMETRIC, force=True)
        tr.set_current_image(g1, d1)
        tr.track()
    tr.synchronize()
    fps = reps / (time.perf_counter() - t0)
    res = {}
    for dt in (np.float32, np.float64):
        o = oracle.Tracker(W, H, cal, mode=oracle.PHOTOMETRIC, levels=3, iters=(10, 10, 10, 0), nselect=8192,
                           mad_bins=256, dtype=dt)
        t1 = time.perf_counter()
        o.set_reference(g0, d0); o.set_current(g1, d1); o.track()
        res[dt] = (o.poses()[0], time.perf_counter() - t1)
    e32 = synth.pose_error(T, res[np.float32][0])
    e64 = synth.pose_error(T, res[np.float64][0])
    et = synth.pose_error(T, Tgt)
    print(f"GPU {fps:.0f} frames/s (host frames in, pose out); vs FP32 oracle {e32[0]:.2e} deg {e32[1]:.2e} mm; "
          f"vs FP64 oracle {e64[0]:.2e} deg {e64[1]:.2e} mm (oracle {res[np.float64][1]:.2f} s on one core); "
          f"vs truth {et[0]:.3f} deg {et[1]:.2f} mm", flush=True)
    assert np.radians(e32[0]) < 1e-5 and e32[1] < 1e-2
    assert e64[0] < 1e-2 and e64[1] < 0.1
    assert et[0] < 0.1 and et[1] < 2.0
