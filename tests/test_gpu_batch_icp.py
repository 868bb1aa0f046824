"""Batched geometric and bi-objective tracking (kb_track with the ICP terms)
against the single-pair tracker and the oracle, on a mixed-size batch
(160x120 .. 640x480).

- The batch's keyframe and current clouds of every level equal the FP32
  oracle's bit for bit.
- The coarsest level's ICP selection count and Student-t variance, formed
  under the initial pose before any update, equal the oracle's bit for bit.
- Every pair's tracked pose agrees with the single-pair tracker and with the
  FP32 oracle tracker, both started from the same pose, within 1e-5 rad /
  1e-2 mm; the error against the FP64 oracle is printed."""
import numpy as np
import pytest

import oracle
from recon_amd import BIOBJECTIVE, GEOMETRIC, synth
from recon_amd.hip import DeviceBuffer
from recon_amd.tracker import BatchTracker, PhotoTracker

pytestmark = pytest.mark.gpu

SIZES = [(160, 120), (320, 240), (480, 360), (640, 480)]
ITERS = (1, 3, 10, 5)
N = 8192


def _pairs(n, seed):
    out = []
    for i in range(n):
        W, H = SIZES[i % 4]
        g0, d0, g1, d1, Tgt = synth.make_pair(seed + i, W, H, 3.0, 50.0)
        init = synth.sample_motion(seed * 5 + i, 1.0, 15.0) @ Tgt
        bufs = [DeviceBuffer.from_array(a) for a in (g0, d0, g1, d1)]
        out.append(dict(ref_gray=bufs[0].ptr, ref_depth=bufs[1].ptr, cur_gray=bufs[2].ptr, cur_depth=bufs[3].ptr,
                        width=W, height=H, calib=synth.calib_block_np(W, H), host=(g0, d0, g1, d1), bufs=bufs,
                        init=init, gt=Tgt))
    return out


@pytest.mark.parametrize("mode", [BIOBJECTIVE, GEOMETRIC])
def test_batch_icp_matches_single_and_oracle(mode):
    print(f"case batch icp mode {mode}", flush=True)
    pairs = _pairs(8, 1200 + 10 * mode)
    bt = BatchTracker(len(pairs), 640, 480, mode=mode, n_select=N, iters=ITERS, lam=0.5)
    Trel, _, rms, st = bt.track(pairs, init=[p["init"] for p in pairs])
    assert np.all(st == 0)
    from tests._oracle_steps import Steps
    for i, p in enumerate(pairs):
        W, H = p["width"], p["height"]
        L = 4 if W == 640 else 3
        g0, d0, g1, d1 = p["host"]
        # clouds of every level, bit for bit
        stp = Steps(W, H, L, p["calib"], g0, d0, g1, d1, nselect=N, dtype=np.float32)
        stp.clouds()
        for l in range(L):
            nl = (W >> l) * (H >> l)
            rc = bt.debug_read(i, "REF_CLOUD", l, np.float32, 8 * nl)
            pos, nrm = rc[:4 * nl].reshape(nl, 4), rc[4 * nl:].reshape(nl, 4)
            R = stp.rpts[l].reshape(nl, 6)
            assert np.array_equal(pos[:, :3], R[:, :3]) and np.array_equal(nrm[:, :3], R[:, 3:]), (i, l)
            assert np.array_equal(pos[:, 3] != 0, stp.rmask[l].astype(bool)), (i, l)
            cp = bt.debug_read(i, "CUR_CLOUD", l, np.float32, 4 * nl).reshape(nl, 4)
            C = stp.cpts[l].reshape(nl, 6)
            assert np.array_equal(cp[:, :3], C[:, :3]) and np.array_equal(cp[:, 3] != 0, stp.cmask[l].astype(bool))
        # the coarsest level's selection under the initial pose, bit for bit
        o = stp.icp_select(L - 1, p["init"].astype(np.float32))
        cnt = bt.debug_read(i, "ICP_COUNT", 0, np.int32, 4)
        var = bt.debug_read(i, "ICP_VAR", L - 1, np.float32, 3)
        assert cnt[L - 1] == o["n"], (i, cnt, o["n"])
        assert var[2] / np.float32(o["n"]) == o["sigma2"], (i, var, o["sigma2"])
        # poses: single-pair tracker and the oracles from the same start
        tr = PhotoTracker(W, H, p["calib"], mode=mode, n_select=N, iters=ITERS, lam=0.5)
        tr.set_reference_points(g0, d0, mode=mode, lam=0.5, force=True)
        tr.set_current_image(g1, d1)
        tr.track_from(np.linalg.inv(p["init"]))
        Ts = tr.trel().astype(np.float64)
        tr.close()
        res = {}
        for dt in (np.float32, np.float64):
            ot = oracle.Tracker(W, H, p["calib"], mode=mode, levels=L, iters=ITERS, nselect=N, lam=0.5, dtype=dt)
            ot.set_reference(g0, d0); ot.set_current(g1, d1); ot.prepare(); ot.set_trel(p["init"]); ot.track()
            res[dt] = ot.poses()[0]
        es = synth.pose_error(Trel[i], Ts)
        e32 = synth.pose_error(Trel[i], res[np.float32])
        e64 = synth.pose_error(Trel[i], res[np.float64])
        et = synth.pose_error(Trel[i], p["gt"])
        print(f"  pair {i} {W}x{H}: vs single {es[0]:.2e} deg {es[1]:.2e} mm; vs FP32 oracle {e32[0]:.2e} deg "
              f"{e32[1]:.2e} mm; vs FP64 oracle {e64[0]:.2e} deg {e64[1]:.2e} mm; vs truth {et[0]:.3f} deg "
              f"{et[1]:.2f} mm; ICP points {list(cnt[:L])}", flush=True)
        assert np.radians(es[0]) < 1e-5 and es[1] < 1e-2, i
        assert np.radians(e32[0]) < 1e-5 and e32[1] < 1e-2, i
    bt.close()
