"""The batched persistent tracker (kb_track) against the oracle at the shape it
is benchmarked on: 640x480, 8192 points per pair, 4 levels, iterations
{1, 3, 10, 5}, where the LDS point cache holds the first 5632 points and the
rest stream past it; plus a mixed-size batch (160x120 .. 640x480) in the same
launch. Every pair's pose is compared with the FP32 oracle tracker (the
reference's arithmetic restated, same initial pose) within 1e-5 rad / 1e-2 mm,
and its error against the FP64 oracle tracker is printed. The comparison runs
for the variant the autotuner picks and for one variant of each family
(residuals in registers or LDS, with or without the point cache), each forced
through RT_BATCH_VARIANT; and the batch preparation equals the oracle bit for
bit at 8192 points."""
import ctypes
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
from recon_amd import PHOTOMETRIC, lib, synth
from recon_amd.hip import DeviceBuffer
from recon_amd.tracker import BatchTracker

pytestmark = pytest.mark.gpu

N = 8192
ITERS = (1, 3, 10, 5)
TOL_RAD, TOL_MM = 1e-5, 1e-2


def _variants():
    """{(rz, with point cache): first variant index of that family covering N points}"""
    fam = {}
    info = (ctypes.c_int * 6)()
    v = 0
    while lib().rt_batch_variant_info(v, info) == 0:
        bs, ppt, rz, pc, _, mode = list(info)
        covers = rz or bs * ppt >= N
        if mode == 1 and covers and (rz, pc > 0) not in fam:
            fam[(rz, pc > 0)] = v
        v += 1
    return fam


def _pairs(sizes, seed, theta=3.0, tmax=50.0):
    """Pairs of the given sizes, warm-started from the true motion perturbed by
    up to 1 deg / 15 mm (what a pose proposal hands Gauss-Newton), so that both
    trackers start inside the basin of the same minimum."""
    out = []
    for i, (W, H) in enumerate(sizes):
        g0, d0, g1, d1, Tgt = synth.make_pair(seed + i, W, H, theta, tmax)
        init = synth.sample_motion(seed * 7 + i, 1.0, 15.0) @ Tgt
        bufs = [DeviceBuffer.from_array(a) for a in (g0, d0, g1, d1)]
        out.append(dict(W=W, H=H, host=(g0, d0, g1, d1), bufs=bufs, gt=Tgt, init=init,
                        calib=synth.calib_block_np(W, H)))
    return out


def _oracle_poses(pairs, dtype):
    def run(p):
        L = 4 if p["W"] == 640 else 3
        o = oracle.Tracker(p["W"], p["H"], p["calib"], mode=oracle.PHOTOMETRIC, levels=L, iters=ITERS, nselect=N,
                           dtype=dtype)
        g0, d0, g1, d1 = p["host"]
        o.set_reference(g0, d0)
        o.set_current(g1, d1)
        o.prepare()
        o.set_trel(p["init"])
        o.track()
        return o.poses()[0]
    with ThreadPoolExecutor(8) as ex:
        return list(ex.map(run, pairs))


def _track(pairs, variant, monkeypatch):
    if variant is None:
        monkeypatch.delenv("RT_BATCH_VARIANT", raising=False)
    else:
        monkeypatch.setenv("RT_BATCH_VARIANT", str(variant))
    bt = BatchTracker(len(pairs), 640, 480, mode=PHOTOMETRIC, n_select=N, iters=ITERS)
    dicts = [dict(ref_gray=p["bufs"][0].ptr, ref_depth=p["bufs"][1].ptr, cur_gray=p["bufs"][2].ptr,
                  cur_depth=p["bufs"][3].ptr, width=p["W"], height=p["H"], calib=p["calib"]) for p in pairs]
    try:
        Trel, _, _, st = bt.track(dicts, init=[p["init"] for p in pairs])
    finally:
        bt.close()
    return Trel, st


CASES = {
    "640x480": [(640, 480)] * 8,
    "mixed": [(160, 120), (320, 240), (480, 360), (640, 480), (640, 480), (480, 360), (320, 240), (160, 120)],
}


@pytest.mark.parametrize("case", list(CASES))
def test_batch_track_matches_oracle(case, monkeypatch):
    print(f"case batch vs oracle {case}, N={N}", flush=True)
    pairs = _pairs(CASES[case], 500 if case == "640x480" else 600)
    o32 = _oracle_poses(pairs, np.float32)
    o64 = _oracle_poses(pairs, np.float64)
    fam = _variants()
    runs = [("autotuned", None)] + [(f"rz{rz} cache{int(pc)} (variant {v})", v) for (rz, pc), v in sorted(fam.items())]
    assert len(runs) == 5, fam
    for name, v in runs:
        print(f"  {name}", flush=True)
        Trel, st = _track(pairs, v, monkeypatch)
        assert np.all(st == 0)
        for i, p in enumerate(pairs):
            r32, t32 = synth.pose_error(Trel[i], o32[i])
            r64, t64 = synth.pose_error(Trel[i], o64[i])
            rt, tt = synth.pose_error(Trel[i], p["gt"])
            print(f"    pair {i} {p['W']}x{p['H']}: vs FP32 oracle {r32:.2e} deg {t32:.2e} mm; "
                  f"vs FP64 oracle {r64:.2e} deg {t64:.2e} mm; vs truth {rt:.3f} deg {tt:.2f} mm", flush=True)
            # the FP64 oracle's error is stated above, not bounded: where the
            # minimum is shallow the FP32 and FP64 trackers settle up to a
            # millimetre apart, and the device follows the FP32 arithmetic
            assert np.radians(r32) < TOL_RAD and t32 < TOL_MM, (name, i)


def test_batch_preparation_bitexact_at_8192_points():
    """Keyframe preparation of a mixed-size batch at 8192 points (the
    benchmark's count) against the FP32 oracle: gradient histogram, selection,
    points, Jacobians and reference gray of every level."""
    from tests._oracle_steps import Steps
    print(f"case batch preparation N={N}", flush=True)
    pairs = _pairs(CASES["mixed"][:4], 700)
    bt = BatchTracker(len(pairs), 640, 480, mode=PHOTOMETRIC, n_select=N, iters=ITERS)
    dicts = [dict(ref_gray=p["bufs"][0].ptr, ref_depth=p["bufs"][1].ptr, cur_gray=p["bufs"][2].ptr,
                  cur_depth=p["bufs"][3].ptr, width=p["W"], height=p["H"], calib=p["calib"]) for p in pairs]
    bt.prepare_only(bt.prepare(dicts))
    color_slot = [14, 17, 20, 23]
    for i, p in enumerate(pairs):
        W, H = p["W"], p["H"]
        L = 4 if W == 640 else 3
        st = Steps(W, H, L, p["calib"], *p["host"], nselect=N, dtype=np.float32)
        st.features()
        assert np.array_equal(bt.debug_read(i, "GRAD_HIST", 0, np.uint32, 256), st.hist), f"histogram pair {i}"
        assert np.array_equal(bt.debug_read(i, "SELECTED", 0, np.int32, N), st.sel), f"selection pair {i}"
        pts = bt.debug_read(i, "POINTS", 0, np.float32, 4 * N).reshape(N, 4)
        rec = st.rec.reshape(-1, 24)[st.sel]
        assert np.array_equal(pts[:, :3], rec[:, :3]), f"points pair {i}"
        for l in range(L):
            jac = bt.debug_read(i, "JACOBIAN", l, np.float32, 7 * N)
            J = st.J[l].reshape(6, N).T
            got = np.concatenate([jac[:4 * N].reshape(N, 4), jac[4 * N:6 * N].reshape(N, 2)], axis=1)
            assert np.array_equal(got, J), f"jacobian pair {i} level {l}"
            assert np.array_equal(jac[6 * N:], rec[:, color_slot[l]]), f"reference gray pair {i} level {l}"
    bt.close()
