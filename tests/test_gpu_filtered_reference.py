"""The filtered reference on the GPU: a keyframe whose ICP cloud comes from the
TSDF model's ray cast (setFilteredReference + setReferencePoints with point
maps), against the oracle: the level-0 cloud bit for bit, the tracked pose
against the FP32 and FP64 oracle trackers fed the same maps, and both against
the true motion; the API's error cases; with the flag off the maps are
ignored."""
import ctypes

import numpy as np
import pytest

import oracle
from recon_amd import BIOBJECTIVE, GEOMETRIC, synth
from recon_amd.hip import DeviceBuffer
from recon_amd.tracker import PhotoTracker
from recon_amd.tsdf import VoxelGrid, default_params

pytestmark = pytest.mark.gpu

W, H, L = 320, 240, 3


def _keyframe_maps(seed):
    """A noise-free pair and the ray cast of the keyframe's depth fused into a
    256^3 grid around the synthetic scene (depths 0.8-3.7 m; the reference's
    default cube ends at 3.1 m), seen from the keyframe pose."""
    g0, d0, g1, d1, Tgt = synth.make_pair(seed, W, H, 2.0, 30.0, noise=0.0)
    fx, fy, cx, cy = synth.pinhole(W, H)
    K = np.array([-fx / W, 0, cx / W, 0, fy / H, cy / H, 0, 0, 1], np.float32)
    vg = VoxelGrid(default_params(origin=(0.0, 0.0, -2000.0), dim=(1700.0, 1700.0, 1700.0)))
    eye = np.eye(4, dtype=np.float32)
    vg.integrate(d0, eye, K)
    xyz, nrm, _ = vg.raycast(eye, K, True)
    return vg, (g0, d0, g1, d1, Tgt), xyz, nrm


@pytest.mark.parametrize("mode", [GEOMETRIC, BIOBJECTIVE])
def test_filtered_reference_matches_oracle(mode):
    print(f"case filtered reference mode {mode}", flush=True)
    vg, (g0, d0, g1, d1, Tgt), xyz, nrm = _keyframe_maps(21)
    assert (xyz[..., 2] < 0).mean() > 0.5
    cal = synth.calib_block_np(W, H)
    tr = PhotoTracker(W, H, cal, mode=mode, levels=L, n_select=8192, iters=(10, 10, 10, 0))
    tr.set_filtered_reference(True)
    assert tr.get_filtered_reference()
    assert tr.set_reference_from_points(vg._xyz.ptr, vg._nrm.ptr, g0, d0, mode=mode, lam=0.5, force=True)
    tr.set_current_image(g1, d1)
    # the level-0 cloud, bit for bit
    tr.debug_prepare()
    n = W * H
    rc = tr.debug_read("REF_CLOUD", 0, np.float32, 8 * n)
    pos, rn = rc[:4 * n].reshape(n, 4), rc[4 * n:].reshape(n, 4)
    pts = np.zeros(n * 6, np.float32)
    mask = np.zeros(n, np.int32)
    oracle.call("or_oriented_cloud_f", np.ascontiguousarray(xyz), np.ascontiguousarray(nrm), W, H, pts, mask)
    m = mask.astype(bool)
    assert np.array_equal(pos[:, 3] != 0, m)
    assert np.array_equal(pos[:, :3], pts.reshape(n, 6)[:, :3])
    assert np.array_equal(rn[:, :3], pts.reshape(n, 6)[:, 3:])
    # tracking
    tr.set_reference_from_points(vg._xyz.ptr, vg._nrm.ptr, g0, d0, mode=mode, lam=0.5, force=True)
    tr.set_current_image(g1, d1)
    assert tr.track()
    Tdev = tr.trel().astype(np.float64)
    for dt in (np.float32, np.float64):
        o = oracle.Tracker(W, H, cal, mode=mode, levels=L, iters=(10, 10, 10, 0), nselect=8192, lam=0.5, dtype=dt)
        o.set_reference_points(g0, d0, xyz, nrm)
        o.set_current(g1, d1)
        o.track()
        rot, tmm = synth.pose_error(Tdev, o.poses()[0])
        print(f"  device vs oracle {dt.__name__}: {rot:.2e} deg, {tmm:.2e} mm", flush=True)
        if dt == np.float32:
            assert np.radians(rot) < 1e-5 and tmm < 1e-2
        else:
            assert rot < 1e-2 and tmm < 0.1
    rot, tmm = synth.pose_error(Tdev, Tgt)
    print(f"  device vs ground truth: {rot:.3f} deg, {tmm:.3f} mm", flush=True)
    assert rot < 0.1 and tmm < 5.0


def test_filtered_reference_api():
    vg, (g0, d0, g1, d1, _), xyz, _ = _keyframe_maps(22)
    cal = synth.calib_block_np(W, H)
    tr = PhotoTracker(W, H, cal, mode=GEOMETRIC, levels=L, n_select=8192)
    assert not tr.get_filtered_reference()
    tr.set_current_image(g1, d1)
    # flag off: the maps are ignored and the cloud is the measured one
    tr.set_reference_from_points(vg._xyz.ptr, vg._nrm.ptr, g0, d0, mode=GEOMETRIC, force=True)
    tr.debug_prepare()
    n = W * H
    a = tr.debug_read("REF_CLOUD", 0, np.float32, 8 * n)
    tr.set_reference_points(g0, d0, mode=GEOMETRIC, force=True)
    tr.debug_prepare()
    b = tr.debug_read("REF_CLOUD", 0, np.float32, 8 * n)
    assert np.array_equal(a, b)
    # flag on: point maps are required, and sequences (measured keyframes) are refused
    tr.set_filtered_reference(True)
    with pytest.raises(RuntimeError):
        tr.set_reference_points(g0, d0, mode=GEOMETRIC, force=True)
    with pytest.raises(RuntimeError):
        tr.set_reference_from_points(None, None, g0, d0, mode=GEOMETRIC, force=True)
    gd, dd = DeviceBuffer(2 * W * H), DeviceBuffer(2 * W * H * 2)
    gd.upload(np.stack([g0, g1])); dd.upload(np.stack([d0, d1]))
    poses = np.zeros(32)
    rc = tr.L.rt_track_sequence(tr.h, gd.ptr, dd.ptr, 2, 10, 0, poses.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc < 0
    tr.set_filtered_reference(False)
    assert tr.set_reference_points(g0, d0, mode=GEOMETRIC, force=True)
