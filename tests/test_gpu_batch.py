"""The batched persistent tracker against the single-pair tracker and the
oracle, on a mixed-size batch (160x120 .. 640x480)."""
import numpy as np
import pytest

from recon_amd import PHOTOMETRIC, synth
from recon_amd.hip import DeviceBuffer
from recon_amd.tracker import BatchTracker, PhotoTracker

pytestmark = pytest.mark.gpu

SIZES = [(160, 120), (320, 240), (480, 360), (640, 480)]


def _batch(n, theta=3.0, tmax=50.0, seed=20, rig=False):
    """n pairs of mixed sizes; rig: every other pair under the two-camera rig
    calibration (distortion, TLR/TRL, KL != KR) instead of the pinhole one."""
    pairs, host, gts = [], [], []
    for i in range(n):
        W, H = SIZES[i % len(SIZES)]
        g0, d0, g1, d1, Tgt = synth.make_pair(seed + i, W, H, theta, tmax)
        bufs = [DeviceBuffer.from_array(a) for a in (g0, d0, g1, d1)]
        host.append((g0, d0, g1, d1, bufs))
        cal = synth.calib_block_rig(W, H) if rig and i % 2 else synth.calib_block_np(W, H)
        pairs.append(dict(ref_gray=bufs[0].ptr, ref_depth=bufs[1].ptr, cur_gray=bufs[2].ptr, cur_depth=bufs[3].ptr,
                          width=W, height=H, calib=cal))
        gts.append(Tgt)
    return pairs, host, gts


@pytest.mark.parametrize("robust,bins,solver,rig", [(0, 64, 0, False), (1, 256, 1, False), (0, 64, 0, True)])
def test_batch_matches_single_pair(robust, bins, solver, rig):
    """Each pair of a mixed-size batch against the single-pair tracker, both
    started from the same pose near the truth (the true motion perturbed by up
    to 1 deg / 15 mm), within 1e-5 rad / 1e-2 mm."""
    print(f"case batch robust {robust} bins {bins} solver {solver} rig {rig}", flush=True)
    n = 8
    N = 2048
    pairs, host, gts = _batch(n, rig=rig)
    init = [synth.sample_motion(900 + i, 1.0, 15.0) @ gts[i] for i in range(n)]
    bt = BatchTracker(n, 640, 480, mode=PHOTOMETRIC, n_select=N, robust=robust, mad_bins=bins, solver=solver)
    Trel, T, rms, st = bt.track(pairs, init=init)
    assert np.all(st == 0)
    for i in range(n):
        W, H = pairs[i]["width"], pairs[i]["height"]
        tr = PhotoTracker(W, H, pairs[i]["calib"], mode=PHOTOMETRIC, n_select=N, robust=robust, mad_bins=bins,
                          solver=solver)
        g0, d0, g1, d1, _ = host[i]
        tr.set_reference_points(g0, d0, mode=PHOTOMETRIC, force=True)
        tr.set_current_image(g1, d1)
        tr.track_from(np.linalg.inv(init[i]))   # the camera pose relative to the keyframe
        Ts = tr.trel().astype(np.float64)
        rot, tmm = synth.pose_error(Trel[i], Ts)
        print(f"  pair {i} {W}x{H}: batch vs single {rot:.2e} deg {tmm:.2e} mm; vs truth {synth.pose_error(Trel[i], gts[i])}",
              flush=True)
        assert np.radians(rot) < 1e-5 and tmm < 1e-2
        assert np.allclose(T[i] @ Trel[i], np.eye(4), atol=1e-5)


def test_batch_retrack_deterministic_and_warm_start():
    print("case batch retrack", flush=True)
    n = 8
    pairs, host, gts = _batch(n, theta=2.0, tmax=30.0, seed=40)
    bt = BatchTracker(n, 640, 480, mode=PHOTOMETRIC, n_select=2048)
    a = bt.track(pairs)[0]
    b = bt.retrack(n=n)[0]
    assert np.array_equal(a, b)
    # the same launch enqueued without waiting, collected by rt_batch_results
    bt.retrack_async()
    r = bt.results(n)[0]
    assert np.array_equal(a, r)
    # starting at the converged pose re-converges next to it (the coarse levels
    # run again, so it may move by a fraction of a millimetre)
    c = bt.retrack(init=a.reshape(n, 16), n=n)[0]
    checked = 0
    for i in range(n):
        r0, t0 = synth.pose_error(a[i], gts[i])
        if r0 > 0.1 or t0 > 3.0:
            continue  # not converged from identity: a restart may still move
        rot, tmm = synth.pose_error(c[i], a[i])
        assert rot < 0.15 and tmm < 3.0
        checked += 1
    assert checked >= 2


@pytest.mark.parametrize("N,rig", [(2048, False), (2048, True)])
def test_batch_preparation_bitexact_against_oracle(N, rig):
    """The batch path's keyframe preparation (gradient bins from LDS-staged
    tiles, histogram, selection, points, Jacobians) equals the FP32 oracle bit
    for bit, pair by pair, on a mixed-size batch (rig: half of the pairs under
    the two-camera rig calibration)."""
    print(f"case batch preparation N={N} rig {rig}", flush=True)
    from tests._oracle_steps import Steps
    n = 4
    pairs, host, gts = _batch(n, seed=60, rig=rig)
    bt = BatchTracker(n, 640, 480, mode=PHOTOMETRIC, n_select=N)
    bt.prepare_only(bt.prepare(pairs))
    for i in range(n):
        W, H = pairs[i]["width"], pairs[i]["height"]
        L = 4 if W == 640 else 3
        g0, d0, g1, d1, _ = host[i]
        st = Steps(W, H, L, pairs[i]["calib"], g0, d0, g1, d1, nselect=N, dtype=np.float32)
        st.features()
        assert np.array_equal(bt.debug_read(i, "GRAD_HIST", 0, np.uint32, 256), st.hist), f"histogram pair {i}"
        assert np.array_equal(bt.debug_read(i, "SELECTED", 0, np.int32, N), st.sel), f"selection pair {i}"
        pts = bt.debug_read(i, "POINTS", 0, np.float32, 4 * N).reshape(N, 4)
        rec = st.rec.reshape(-1, 24)[st.sel]
        assert np.array_equal(pts[:, :3], rec[:, :3]), f"points pair {i}"
        color_slot = [14, 17, 20, 23]
        for l in range(L):
            jac = bt.debug_read(i, "JACOBIAN", l, np.float32, 7 * N)
            J = st.J[l].reshape(6, N).T
            got = np.concatenate([jac[:4 * N].reshape(N, 4), jac[4 * N:6 * N].reshape(N, 2)], axis=1)
            assert np.array_equal(got, J), f"jacobian pair {i} level {l}"
            assert np.array_equal(jac[6 * N:], rec[:, color_slot[l]]), f"reference gray pair {i} level {l}"


@pytest.mark.parametrize("N,iters", [(8192, (1, 3, 10, 5)), (12288, (0, 3, 4, 2))])
def test_batch_kernel_variants_agree(N, iters, monkeypatch, capfd):
    """Every variant of the persistent kernel the autotuner may pick, forced one
    at a time: a thread sums its points in the same order (t, t + BS, ...)
    whatever its unroll depth and wherever the residuals or the point records
    are held (registers, LDS residuals, the LDS point cache), so variants of
    one workgroup size agree bit for bit. N = 12288 runs past the point
    cache's 8192 points, and a first level without iterations leaves the
    cache to be filled by the next level."""
    print(f"case batch variants N {N} iters {iters}", flush=True)
    pairs, host, _ = _batch(6, theta=2.0, tmax=30.0, seed=60)
    pairs = [p for p in pairs if p["width"] * p["height"] >= 2 * N]  # at most half the pixels are selected
    monkeypatch.setenv("RT_VERBOSE", "1")
    by_bs, kinds = {}, set()
    for v in range(40):
        monkeypatch.setenv("RT_BATCH_VARIANT", str(v))
        bt = BatchTracker(len(pairs), 640, 480, mode=PHOTOMETRIC, n_select=N, iters=iters)
        capfd.readouterr()
        try:
            Trel, _, rms, st = bt.track(pairs)
        except RuntimeError:
            continue  # no such variant, or it does not cover N points
        finally:
            bt.close()
        line = [s for s in capfd.readouterr().err.splitlines() if "rt_batch autotune" in s][-1]
        f = line.split("->")[1].split()
        cfg = dict(zip(f[0:10:2], map(int, f[1:10:2])))
        print(f"  variant {v}: {cfg}", flush=True)
        assert np.all(st == 0)
        kinds.add((cfg["rz"], cfg["pc"] > 0))
        ref = by_bs.setdefault(cfg["bs"], (v, Trel, rms))
        assert np.array_equal(Trel, ref[1]) and np.array_equal(rms, ref[2]), f"variant {v} vs {ref[0]}"
    assert (0, True) in kinds and (0, False) in kinds
    if N <= 8192:
        assert (1, True) in kinds and (1, False) in kinds


def test_depth_blocks_are_the_current_depth_map():
    """The tracker's depth taps read kb_pyramid's copy of the current depth
    map in 8x8-pixel blocks: every pixel (x, y) must hold the input's value at
    ((y // 8) * (W // 8) + x // 8) * 64 + (y % 8) * 8 + x % 8, also for a
    height that is not a multiple of 8 (236 = 29 * 8 + 4, three levels)."""
    print("case depth blocks 640x480 and 320x236", flush=True)
    sizes = [(640, 480), (320, 236), (640, 480)]
    pairs, host = [], []
    for i, (W, H) in enumerate(sizes):
        g0, d0, g1, d1, _ = synth.make_pair(60 + i, W, H, 3.0, 50.0)
        bufs = [DeviceBuffer.from_array(a) for a in (g0, d0, g1, d1)]
        host.append((d1, bufs))
        pairs.append(dict(ref_gray=bufs[0].ptr, ref_depth=bufs[1].ptr, cur_gray=bufs[2].ptr, cur_depth=bufs[3].ptr,
                          width=W, height=H, calib=synth.calib_block_np(W, H)))
    bt = BatchTracker(len(pairs), 640, 480, mode=PHOTOMETRIC, n_select=2048, iters=(1, 1, 1, 1))
    bt.prepare_only(bt.prepare(pairs))
    for i, (W, H) in enumerate(sizes):
        d1 = np.asarray(host[i][0]).reshape(H, W)
        H8 = (H + 7) // 8 * 8
        blk = bt.debug_read(i, "DEPTH_BLOCKS", 0, np.uint16, W * H8)
        y, x = np.mgrid[0:H, 0:W]
        idx = ((y // 8) * (W // 8) + x // 8) * 64 + (y % 8) * 8 + x % 8
        assert np.array_equal(blk[idx], d1), f"pair {i} ({W}x{H})"
