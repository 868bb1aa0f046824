"""TSDF fusion and ray casting on the GPU against the oracle, bit for bit:
grids after each fused frame (16-bit depth and the reference's float4 RGBD
input), ray-cast points, normals and shading from fused and unfused poses,
with and without the zero-crossing refinement, grid resampling; at a small
grid and at the reference's own size (256^3 voxels, 320x240). The fused
scene is found on its analytic surfaces."""
import numpy as np
import pytest

from oracle.tsdf import Grid
from recon_amd.tsdf import VoxelGrid, default_params
from _tsdf_scene import normalized_k, pose, render_depth, surface_distance, to_world

pytestmark = pytest.mark.gpu

ORIGIN = (0.0, 0.01, -1600.01)
DIM = (1500.0, 1500.0, 1500.0)
POSES = [pose(), pose(2, -4, (80, -20, 30)), pose(-3, 3, (-60, 40, -50))]
UNFUSED = pose(1, 1, (20, 10, 0))


def _bits_equal(a, b):
    a, b = np.ascontiguousarray(a, np.float32), np.ascontiguousarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def _grids(res, W, H):
    print(f"tsdf grid res={res} image={W}x{H}", flush=True)
    g = VoxelGrid(default_params(resolution=res, width=W, height=H))
    o = Grid(res, W, H, ORIGIN, DIM)
    return g, o


@pytest.mark.parametrize("res,W,H", [(64, 160, 120), (128, 160, 120), (256, 320, 240)])
def test_update_and_raycast_match_oracle_bitexact(res, W, H):
    K = normalized_k(W, H)
    g, o = _grids(res, W, H)
    assert _bits_equal(g.read(), o.data)
    for i, T in enumerate(POSES):
        d = np.round(render_depth(T, K, W, H)).astype(np.uint16)
        g.integrate(d, T, K)
        o.update(d.astype(np.float32), T, K)
        assert _bits_equal(g.read(), o.data), f"grid after frame {i}"
        for refine in (True, False):
            for Tr in (T, UNFUSED):
                got = g.raycast(Tr, K, refine)
                exp = o.raycast(Tr, K, refine)
                for name, a, b in zip(("xyz", "normals", "shade"), got, exp):
                    assert _bits_equal(a, b), f"{name} after frame {i}, refine={refine}"
    # the fused scene is where it should be
    xyz, _, _ = g.raycast(UNFUSED, K, True)
    v = xyz[..., 3] > 0
    v[:4] = v[-4:] = False
    v[:, :4] = v[:, -4:] = False
    assert v.mean() > 0.6
    assert np.median(surface_distance(to_world(UNFUSED, xyz))[v]) < 1.0


def test_rgbd_input_matches_oracle_bitexact():
    W, H, res = 160, 120, 64
    K = normalized_k(W, H)
    g, o = _grids(res, W, H)
    mn, mx = 300.0, 7000.0
    for T in POSES[:2]:
        z = render_depth(T, K, W, H).astype(np.float32)
        rgbd = np.zeros((H, W, 4), np.float32)
        rgbd[..., :3] = 0.5
        rgbd[..., 3] = np.where(z > 0, (z - mn) / (mx - mn), 0)[::-1]  # rows bottom-up
        g.set_depth(rgbd)
        g.update(T, K)
        o.update(o.dist_from_rgbd(rgbd), T, K)
        assert _bits_equal(g.read(), o.data)
    assert (g.read()[..., 1] > 0).sum() > 10000


def test_resample_matches_oracle_bitexact():
    W, H, res = 160, 120, 64
    K = normalized_k(W, H)
    src, osrc = _grids(res, W, H)
    for T in POSES:
        d = np.round(render_depth(T, K, W, H)).astype(np.uint16)
        src.integrate(d, T, K)
        osrc.update(d.astype(np.float32), T, K)
    dst, odst = _grids(res, W, H)
    for T in (POSES[1], pose(5, -7, (100, 50, -80))):
        dst.resample_from(src, T)
        odst.resample_from(osrc, T)
        assert _bits_equal(dst.read(), odst.data)
    # the ray cast of a resampled grid (its occupancy is rebuilt by the resampling)
    for refine in (True, False):
        got = dst.raycast(UNFUSED, K, refine)
        exp = odst.raycast(UNFUSED, K, refine)
        for name, a, b in zip(("xyz", "normals", "shade"), got, exp):
            assert _bits_equal(a, b), f"{name} of the resampled grid, refine={refine}"
    # the resampled grid, ray cast from the matching pose, sees the same scene
    T = POSES[1]
    dst.resample_from(src, T)
    a, _, _ = dst.raycast(np.eye(4, dtype=np.float32), K, True)
    b, _, _ = src.raycast(T, K, True)
    v = (a[..., 3] > 0) & (b[..., 3] > 0)
    assert v.mean() > 0.5
    assert np.median(np.abs(a[..., 2] - b[..., 2])[v]) < 2.0


def test_reset_and_argument_errors():
    W, H, res = 96, 72, 32
    K = normalized_k(W, H)
    g, _ = _grids(res, W, H)
    g.integrate(np.full((H, W), 1500, np.uint16), np.eye(4, dtype=np.float32), K)
    assert (g.read()[..., 1] > 0).any()
    g.reset()
    r = g.read()
    assert np.all(r[..., 0] == 1) and np.all(r[..., 1] == 0)
    xyz, _, _ = g.raycast(np.eye(4, dtype=np.float32), K, True)
    assert np.all(xyz == 0)
    other = VoxelGrid(default_params(resolution=64, width=W, height=H))
    with pytest.raises(RuntimeError):
        other.resample_from(g, np.eye(4))
    with pytest.raises(RuntimeError):
        g.resample_from(g, np.eye(4))
    with pytest.raises(ValueError):
        g.set_depth(np.zeros((H + 1, W), np.uint16))
