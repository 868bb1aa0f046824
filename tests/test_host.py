"""Host-side checks that need no GPU: the C ABI library loads and exports its
header, parameter and calibration helpers, argument validation, the seeded
synthetic generator, and the oracle tracker end to end."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

import oracle
from recon_amd import RtParams, lib, synth
from recon_amd.tracker import default_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    names = set()
    for h in sorted(glob.glob(os.path.join(ROOT, "include", "rt_*.h"))):
        hdr = open(h).read()
        hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
        names |= set(re.findall(r"\b(rt_[a-z0-9_]+)\s*\(", hdr))
    assert len(names) >= 55
    L = lib()
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_default_params_are_the_reference_defaults():
    p = default_params()
    # phototracker.cpp:134-135 (mode, lambda), :148-149 (iterations), :48 (points), :127-128 (thresholds)
    assert p.mode == 0 and abs(p.lambda_ - 0.9) < 1e-7
    assert list(p.iters) == [1, 3, 10, 5] and p.n_select == 8192 and p.levels == 0
    assert p.mad_bins == 64 and p.robust == 0
    assert p.pose_dist_threshold == 125.0 and p.pose_angle_threshold == 12.5


def test_calib_block_layout():
    c = synth.calib_block(640, 480)
    n = synth.calib_block_np(640, 480)
    assert np.array_equal(c, n)
    # calib.h offsets: KR 0, iKR 9, kcR 18, TLR 26, KL 43, iKL 52, min 62, max 63, TRL 64
    assert c[0] == -525 and c[4] == 525 and c[2] == 319.5 and c[5] == 239.5
    assert np.array_equal(c[43:52], c[0:9])
    K = c[0:9].reshape(3, 3).astype(np.float64)
    assert np.allclose(K @ c[9:18].reshape(3, 3), np.eye(3), atol=1e-4)  # an FP32 inverse
    assert np.array_equal(c[26:42].reshape(4, 4), np.eye(4)) and np.array_equal(c[64:80].reshape(4, 4), np.eye(4))
    assert c[62] == 300 and c[63] == 7000 and np.all(c[18:26] == 0)
    bad = (ctypes.c_float * 96)()
    assert lib().rt_calib_pinhole(-1.0, 1.0, 0, 0, 300, 7000, bad) < 0


@pytest.mark.parametrize("kw,W,H", [
    (dict(n_select=100), 320, 240),           # not a multiple of 256
    (dict(n_select=1 << 20), 320, 240),       # more than half the pixels
    (dict(mad_bins=32), 320, 240),
    (dict(levels=4), 100, 60),                # not divisible by 2^(levels-1)
    (dict(), 2048, 1024),                     # more than 1024 pixels per selection segment
    (dict(mode=7), 320, 240),
])
def test_create_rejects_bad_parameters_before_touching_the_device(kw, W, H):
    p = default_params(**kw)
    h = ctypes.c_void_p()
    cal = synth.calib_block_np(W, H)
    rc = lib().rt_create(0, W, H, cal.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ctypes.byref(p), ctypes.byref(h))
    assert rc == -1 and not h.value
    assert lib().rt_last_error()


def test_null_arguments_are_errors():
    L = lib()
    assert L.rt_track(None) == -1
    assert L.rt_set_current(None, None, None, 0) == -1
    assert L.rt_get_global_pose(None, None) == -1
    assert L.rt_batch_track(None, 1, None, None) == -1


def test_synthetic_frames_are_seeded_and_in_range():
    g0, d0, g1, d1, Tgt = synth.make_pair(3, 160, 120, 2.0, 30.0)
    a0, b0, a1, b1, T2 = synth.make_pair(3, 160, 120, 2.0, 30.0)
    assert np.array_equal(g0, a0) and np.array_equal(d1, b1) and np.array_equal(Tgt, T2)
    assert d0.min() >= 500 and d0.max() <= 4000 and (d0 > 0).mean() > 0.95
    assert 20 < g0.std() < 120
    rot, tr = synth.pose_error(np.eye(4), Tgt)
    assert rot <= 2.0 + 1e-9 and tr <= 30.0 + 1e-9


def test_motion_sampler_bounds():
    for s in range(200):
        M = synth.sample_motion(s, 3.0, 50.0)
        rot, tr = synth.pose_error(M, np.eye(4))
        assert rot <= 3.0 + 1e-9 and tr <= 50.0 + 1e-9
        assert np.allclose(M[:3, :3] @ M[:3, :3].T, np.eye(3), atol=1e-12)


def test_trajectory_increments_are_bounded():
    P = synth.trajectory(5, 300, 1.5, 20.0)
    assert np.array_equal(P[0], np.eye(4))
    for k in range(1, 300):
        rot, tr = synth.pose_error(P[k], P[k - 1])
        assert rot <= 1.5 + 1e-6 and tr <= 20.0 + 1e-6
    # stays inside the room
    assert np.abs(P[:, :3, 3]).max() < 600


def test_oracle_tracker_converges_to_ground_truth():
    """End to end on the CPU: the FP64 oracle tracks a 160x120 pair from
    identity to the true motion in every mode."""
    W, H = 160, 120
    g0, d0, g1, d1, Tgt = synth.make_pair(9, W, H, 2.0, 20.0)
    cal = synth.calib_block_np(W, H)
    start = synth.pose_error(np.eye(4), Tgt)
    for mode in (oracle.PHOTOMETRIC, oracle.GEOMETRIC, oracle.BIOBJECTIVE):
        o = oracle.Tracker(W, H, cal, mode=mode, levels=3, iters=(10, 10, 10, 0), nselect=2048, lam=0.5)
        o.set_reference(g0, d0); o.set_current(g1, d1); o.track()
        Trel, T, Tinc, G = o.poses()
        rot, tr = synth.pose_error(Trel, Tgt)
        assert rot < 0.25 * start[0] and tr < 0.25 * start[1], (mode, rot, tr, start)
        assert np.allclose(T @ Trel, np.eye(4), atol=1e-5) and np.allclose(G, T, atol=1e-4)


def test_oracle_fp32_and_fp64_trackers_agree():
    W, H = 160, 120
    g0, d0, g1, d1, Tgt = synth.make_pair(10, W, H, 1.5, 15.0)
    cal = synth.calib_block_np(W, H)
    out = []
    for dt in (np.float32, np.float64):
        o = oracle.Tracker(W, H, cal, mode=oracle.PHOTOMETRIC, levels=3, iters=(10, 10, 10, 0), nselect=2048, dtype=dt)
        o.set_reference(g0, d0); o.set_current(g1, d1); o.track()
        out.append(o.poses()[0])
    rot, tr = synth.pose_error(out[0], out[1])
    assert rot < 1e-2 and tr < 0.1


def test_cpp_phototracker_header_compiles_and_links(tmp_path):
    """The PhotoTracker-compatible C++ class builds against the C ABI with a
    plain host compiler and reports a rejected configuration the reference's
    way (message on stderr, null tracker) without touching a device."""
    import subprocess
    src = tmp_path / "pt.cpp"
    src.write_text(
        '#include "rt_phototracker.hpp"\n'
        "int main() {\n"
        "    float cal[RT_CALIB_SIZE];\n"
        "    if (rt_calib_pinhole(525, 525, 319.5f, 239.5f, 300, 7000, cal) != RT_OK) return 1;\n"
        "    rt::PhotoTracker* t = rt::PhotoTracker::create(0, 2048, 1024, cal);\n"
        "    return t == nullptr ? 0 : 2;\n"
        "}\n")
    exe = tmp_path / "pt"
    libdir = os.path.join(ROOT, "recon_amd")
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src),
                    "-L", libdir, "-lrecon_amd", f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0
    assert "PhotoTracker::initializePhotoTracker" in r.stderr
