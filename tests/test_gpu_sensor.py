"""The sensor-side inputs on the GPU: the disparity -> depth kernel against the
oracle bit for bit (synthetic disparities with holes, and real Kinect frames
of the reference's sample sequence), the 16-bit depth rounding, and the
tracker on those real frames (calibration from the reference's calib.xml,
two-camera rig with distortion): frame-to-frame motions stay small and
tracking a pair in both directions gives inverse poses."""
import os

import numpy as np
import pytest

from oracle import sensor as osn
from recon_amd import BIOBJECTIVE, PHOTOMETRIC, sensor
from recon_amd.hip import DeviceBuffer, hip
from recon_amd.tracker import PhotoTracker

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CALIB_XML = os.path.join(HERE, "golden", "calib.xml")


def _frames():
    return np.load(os.path.join(HERE, "golden", "kinect_frames.npz"))


@pytest.mark.parametrize("half", [True, False])
@pytest.mark.parametrize("xy_offset", [True, False])
@pytest.mark.parametrize("mode", [0, 1])
def test_disparity_to_depth_matches_oracle_bitexact(half, xy_offset, mode):
    print(f"case disparity half {half} offset {xy_offset} mode {mode}", flush=True)
    c, _ = sensor.load_calib_xml(CALIB_XML, 320 if half else 640, 240 if half else 480)
    rng = np.random.default_rng(11)
    d = rng.integers(350, 1090, (480, 640)).astype(np.uint16)  # up to beyond the 7 m window
    d[rng.random(d.shape) < 0.08] = 2047
    d[100:140, 200:260] = 2047  # a hole
    got = sensor.disparity_to_depth(d, c, xy_offset=xy_offset, half=half, mode=mode)
    want = osn.disparity_to_depth_np(d, c, xy_offset, half, mode)
    assert got.shape == want.shape
    assert np.array_equal(got, want), np.argwhere(got != want)[:5]
    # odd sizes: the unwritten tail rows and the borders
    d2 = d[:94, :126]
    got2 = sensor.disparity_to_depth(d2, c, xy_offset=xy_offset, half=half, mode=mode)
    assert np.array_equal(got2, osn.disparity_to_depth_np(d2, c, xy_offset, half, mode))


def test_disparity_to_depth_real_frames_bitexact():
    print("case disparity real frames", flush=True)
    f = _frames()
    c, off = sensor.load_calib_xml(CALIB_XML, 320, 240)
    for k in range(1, 5):
        d = f[f"disparity{k}"]
        for mode in (0, 1):
            got = sensor.disparity_to_depth(d, c, xy_offset=off, half=True, mode=mode)
            assert np.array_equal(got, osn.disparity_to_depth_np(d, c, off, True, mode)), (k, mode)
        z = sensor.disparity_to_depth(d, c, xy_offset=off, half=True)
        assert 0.5 < (z > 0).mean() < 1.0 and 300 <= z[z > 0].min() and z.max() <= 7000


def test_depth_mm_to_u16():
    print("case depth rounding", flush=True)
    z = np.array([0.0, -5.0, 0.49, 0.5, 1.5, 1234.49, 1234.5, 65535.4, 65535.6, 7000.0], np.float32)
    zb = DeviceBuffer.from_array(z)
    out = DeviceBuffer(2 * len(z))
    sensor.depth_mm_to_u16(zb, len(z), out)
    hip().hipDeviceSynchronize()
    got = out.to_array(np.uint16, len(z))
    assert list(got) == [0, 0, 0, 1, 2, 1234, 1235, 65535, 0, 7000]


def _track(W, H, c, ref, cur, mode=BIOBJECTIVE):
    tr = PhotoTracker(W, H, c, mode=mode, n_select=8192, lambda_=0.5)
    tr.set_reference_points(ref[0], ref[1], mode=mode, lam=0.5, force=True)
    tr.set_current_image(cur[0], cur[1])
    tr.track_from(np.eye(4))
    T = tr.get_relative_pose().astype(np.float64)
    rms = tr.residual_rms()
    tr.close()
    return T, rms


def _rot_deg(T):
    return float(np.degrees(np.arccos(np.clip((np.trace(T[:3, :3]) - 1) / 2, -1, 1))))


@pytest.mark.parametrize("mode", [PHOTOMETRIC, BIOBJECTIVE])
def test_tracks_real_kinect_frames(mode):
    """Frames 1-4 of the reference's sample sequence at 320x240: depth from
    the disparity kernel (IR grid, half resolution), gray from the Bayer
    mosaic, calibration from calib.xml (distortion, TLR). Consecutive frames of
    a hand-held 30 Hz capture move little, and tracking a pair in both
    directions gives inverse poses."""
    print(f"case real kinect mode {mode}", flush=True)
    f = _frames()
    W, H = 320, 240
    c, off = sensor.load_calib_xml(CALIB_XML, W, H)
    frames = []
    for k in range(1, 5):
        z = sensor.disparity_to_depth(f[f"disparity{k}"], c, xy_offset=off, half=True)
        d16 = np.where(z > 0, np.rint(z), 0).astype(np.uint16)
        frames.append((sensor.bayer_to_gray_half(f[f"bayer{k}"]), d16))
    for a, b in [(0, 1), (1, 2), (2, 3), (0, 2)]:
        Tab, rab = _track(W, H, c, frames[a], frames[b], mode)
        Tba, rba = _track(W, H, c, frames[b], frames[a], mode)
        E = Tab @ Tba
        print(f"  frames {a + 1}->{b + 1}: {_rot_deg(Tab):.3f} deg {np.linalg.norm(Tab[:3, 3]):.2f} mm, rms {rab:.4f}; "
              f"round trip {_rot_deg(E):.4f} deg {np.linalg.norm(E[:3, 3]):.3f} mm", flush=True)
        assert np.isfinite(rab) and np.isfinite(rba)
        assert _rot_deg(Tab) < 5.0 and np.linalg.norm(Tab[:3, 3]) < 100.0
        assert _rot_deg(E) < 0.5 and np.linalg.norm(E[:3, 3]) < 10.0
