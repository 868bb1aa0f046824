"""Degenerate inputs on the GPU: frames without valid depth, a system that
cannot be solved, and geometric levels without points. The reference has no
tests; its kernels handle these by construction (invalid points carry
z = -maxDist and zero gradients, d2z.cu:330; the residual outside the view is
1 with weight 0, warp.cu:298-303; an empty ICP selection leaves no system,
shapetracker.cpp:357-382). Here every path must end normally: no update and
a failed status where the Cholesky solve meets a non-positive pivot, the
initial pose kept where no geometric point exists."""
import numpy as np
import pytest

from recon_amd import BIOBJECTIVE, GEOMETRIC, PHOTOMETRIC, synth
from recon_amd.hip import DeviceBuffer
from recon_amd.tracker import BatchTracker, PhotoTracker

pytestmark = pytest.mark.gpu

W, H, N = 320, 240, 2048


def _pair(seed, kill_ref_depth=False, kill_cur_depth=False):
    g0, d0, g1, d1, Tgt = synth.make_pair(seed, W, H, 2.0, 30.0)
    if kill_ref_depth:
        d0 = np.zeros_like(d0)
    if kill_cur_depth:
        d1 = np.zeros_like(d1)
    bufs = [DeviceBuffer.from_array(a) for a in (g0, d0, g1, d1)]
    return dict(ref_gray=bufs[0].ptr, ref_depth=bufs[1].ptr, cur_gray=bufs[2].ptr, cur_depth=bufs[3].ptr, width=W,
                height=H, calib=synth.calib_block_np(W, H), host=(g0, d0, g1, d1), bufs=bufs, gt=Tgt)


def test_keyframe_without_valid_depth():
    """A keyframe whose depth is all invalid: every selected point is empty, the
    photometric system is zero, the Cholesky solve fails: the batch reports
    status -1 and keeps the initial pose; the single-pair tracker keeps its
    pose; a valid pair in the same batch is unaffected."""
    print("case keyframe without depth", flush=True)
    bad, good = _pair(1, kill_ref_depth=True), _pair(2)
    init = np.eye(4)
    bt = BatchTracker(2, W, H, mode=PHOTOMETRIC, n_select=N)
    Trel, _, rms, st = bt.track([bad, good], init=[init, init])
    assert st[0] == -1 and st[1] == 0
    assert np.array_equal(Trel[0], init)
    bt1 = BatchTracker(1, W, H, mode=PHOTOMETRIC, n_select=N)
    alone, _, _, st1 = bt1.track([good], init=[init])
    r, t = synth.pose_error(Trel[1], alone[0])
    assert st1[0] == 0 and np.radians(r) < 1e-5 and t < 1e-2
    tr = PhotoTracker(W, H, bad["calib"], mode=PHOTOMETRIC, n_select=N)
    g0, d0, g1, d1 = bad["host"]
    tr.set_reference_points(g0, d0, mode=PHOTOMETRIC, force=True)
    tr.set_current_image(g1, d1)
    assert tr.track()
    assert np.array_equal(tr.trel(), np.eye(4, dtype=np.float32))


def test_unsolvable_start_pose():
    """A start pose of NaNs projects nothing into the view: the system is zero,
    the solve fails, the status is -1 and the launch ends normally; the next
    pair of the batch is tracked as usual."""
    print("case NaN start pose", flush=True)
    p, q = _pair(3), _pair(4)
    nan = np.full((4, 4), np.nan)
    bt = BatchTracker(2, W, H, mode=PHOTOMETRIC, n_select=N)
    Trel, _, rms, st = bt.track([p, q], init=[nan, np.eye(4)])
    assert st[0] == -1 and st[1] == 0
    bt1 = BatchTracker(1, W, H, mode=PHOTOMETRIC, n_select=N)
    alone, _, _, st1 = bt1.track([q], init=[np.eye(4)])
    r, t = synth.pose_error(Trel[1], alone[0])
    assert st1[0] == 0 and np.radians(r) < 1e-5 and t < 1e-2


@pytest.mark.parametrize("mode", [GEOMETRIC, BIOBJECTIVE])
def test_geometric_levels_without_points(mode):
    """A current frame without valid depth leaves every ICP level without
    points: no update is made (as the single-pair path, which skips the solve
    of an empty selection), the initial pose stays, the counts are zero."""
    print(f"case empty ICP selection mode {mode}", flush=True)
    p = _pair(5, kill_cur_depth=True)
    init = synth.sample_motion(77, 1.0, 10.0)
    bt = BatchTracker(1, W, H, mode=mode, n_select=N, lam=0.5)
    Trel, _, rms, st = bt.track([p], init=[init])
    assert st[0] == 0
    assert np.allclose(Trel[0], init.astype(np.float32), atol=0)
    assert np.all(bt.debug_read(0, "ICP_COUNT", 0, np.int32, 4)[:3] == 0)
    tr = PhotoTracker(W, H, p["calib"], mode=mode, n_select=N, lam=0.5)
    g0, d0, g1, d1 = p["host"]
    tr.set_reference_points(g0, d0, mode=mode, lam=0.5, force=True)
    tr.set_current_image(g1, d1)
    tr.track_from(np.linalg.inv(init))
    assert np.allclose(tr.trel().astype(np.float64), init.astype(np.float32), atol=1e-6)
    assert np.all(tr.debug_read("ICP_COUNT", 0, np.int32, 4)[:3] == 0)
