"""rt_batch_track_noalloc (batched tracking whose working memory is the
library's fixed device scratch array, for callers that cannot allocate)
against rt_batch_track on the same pairs, through the C ABI: two calls in a
row, the second with fewer and smaller pairs, so that the second call runs on
scratch the first one left behind. Both run the same tracking kernel variant
(forced through RT_BATCH_VARIANT for the batch object), so the poses agree bit
for bit; the bound asserted is the parity bar, 1e-5 rad / 1e-2 mm."""
import ctypes

import numpy as np
import pytest

from recon_amd import PHOTOMETRIC, RtPair, RtPairResult, lib, synth
from recon_amd.hip import DeviceBuffer
from recon_amd.tracker import BatchTracker, _results, default_params

pytestmark = pytest.mark.gpu

N = 8192
ITERS = (1, 3, 10, 5)


def _pairs(sizes, seed):
    out = []
    for i, (W, H) in enumerate(sizes):
        g0, d0, g1, d1, Tgt = synth.make_pair(seed + i, W, H, 3.0, 50.0)
        init = synth.sample_motion(seed * 3 + i, 1.0, 15.0) @ Tgt
        bufs = [DeviceBuffer.from_array(a) for a in (g0, d0, g1, d1)]
        out.append(dict(ref_gray=bufs[0].ptr, ref_depth=bufs[1].ptr, cur_gray=bufs[2].ptr, cur_depth=bufs[3].ptr,
                        width=W, height=H, calib=synth.calib_block_np(W, H), init=init, bufs=bufs, gt=Tgt))
    return out


def _noalloc(pairs, params):
    n = len(pairs)
    arr = (RtPair * n)()
    keep = []
    for i, p in enumerate(pairs):
        c = np.ascontiguousarray(p["calib"], np.float32)
        t = np.ascontiguousarray(p["init"], np.float64).ravel()
        keep += [c, t]
        arr[i].ref_gray, arr[i].ref_depth = p["ref_gray"], p["ref_depth"]
        arr[i].cur_gray, arr[i].cur_depth = p["cur_gray"], p["cur_depth"]
        arr[i].width, arr[i].height = p["width"], p["height"]
        arr[i].calib = c.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        arr[i].init_T = t.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    out = (RtPairResult * n)()
    rc = lib().rt_batch_track_noalloc(0, n, arr, ctypes.byref(params), None, out)
    return rc, out


def _fixed_variant():
    info = (ctypes.c_int * 6)()
    v = 0
    while lib().rt_batch_variant_info(v, info) == 0:
        bs, ppt, rz, pc, _, mode = list(info)
        if mode == 1 and bs == 1024 and rz == 1 and pc > 0 and ppt == 4:
            return v
        v += 1
    raise AssertionError("no 1024-thread variant with LDS residuals and point cache")


def test_noalloc_matches_batch_two_calls(monkeypatch):
    params = default_params(mode=PHOTOMETRIC, n_select=N, iters=ITERS)
    monkeypatch.setenv("RT_BATCH_VARIANT", str(_fixed_variant()))
    calls = [[(640, 480), (480, 360), (320, 240), (160, 120), (640, 480), (320, 240)],
             [(320, 240), (160, 120), (160, 120)]]
    for k, sizes in enumerate(calls):
        print(f"case no-alloc call {k}: {len(sizes)} pairs {sizes}", flush=True)
        pairs = _pairs(sizes, 800 + 50 * k)
        rc, out = _noalloc(pairs, params)
        assert rc == 0, lib().rt_last_error()
        Tn, _, rms_n, st_n = _results(out)
        bt = BatchTracker(len(pairs), 640, 480, mode=PHOTOMETRIC, n_select=N, iters=ITERS)
        Tb, _, rms_b, st_b = bt.track(pairs, init=[p["init"] for p in pairs])
        bt.close()
        assert np.array_equal(st_n, st_b) and np.all(st_n == 0)
        worst = (0.0, 0.0)
        for i in range(len(pairs)):
            r, t = synth.pose_error(Tn[i], Tb[i])
            worst = (max(worst[0], r), max(worst[1], t))
            assert np.radians(r) < 1e-5 and t < 1e-2, (k, i)
        print(f"  no-alloc vs batch: max {worst[0]:.2e} deg {worst[1]:.2e} mm; bit-equal {np.array_equal(Tn, Tb)}",
              flush=True)
        assert np.array_equal(Tn, Tb) and np.array_equal(rms_n, rms_b)


def test_noalloc_rejects_sizes_beyond_its_limits():
    print("case no-alloc limits", flush=True)
    pairs = _pairs([(160, 120)], 990)
    too_many = pairs * 9
    rc, _ = _noalloc(too_many, default_params(mode=PHOTOMETRIC, n_select=2048))
    assert rc == -1 and b"RT_NOALLOC_MAX_PAIRS" in lib().rt_last_error()
    rc, _ = _noalloc(pairs, default_params(mode=PHOTOMETRIC, n_select=2 * N))
    assert rc == -1 and b"RT_NOALLOC_MAX_POINTS" in lib().rt_last_error()
