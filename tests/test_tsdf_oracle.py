"""The TSDF oracle (oracle/tsdf.hpp) pinned by closed forms: the reference has
no tests or stored outputs for its voxel grid (reconstruct/VoxelGrid.cpp,
kernels/VoxelGrid_cuda.cu), so each function is checked against what the
geometry fixes.

- update: a fronto-parallel plane gives every voxel clamp((Z - depth) /
  (0.03 dim), -1, 1) where the voxel projects inside the image and is not
  beyond the -1.25 truncation; the weight is 1 and saturates at 16.
- raycast: the plane is found at z = -Z with normal (0, 0, 1); a grid holding
  a linear field (trilinear interpolation reproduces it exactly) is cut at
  the field's zero plane, with that plane's normal, from a rotated pose; a
  sphere in front of a tilted wall, fused from three poses, is found on the
  analytic surfaces, pixel by pixel at the analytic depth, from those poses
  and from a fourth one (a wrong pose convention, inverse or transposed
  rotation, misses by centimetres).
- resample: trilinear interpolation reproduces a linear field exactly, so a
  grid holding a linear function, resampled under a rigid pose, holds that
  function at the mapped voxel centres.
- the RGBD -> distance map conversion: the closed form per pixel, rows flipped.
"""
import ctypes

import numpy as np
import pytest

from oracle.tsdf import Grid
from _tsdf_scene import normalized_k, pose, render_depth, surface_distance, to_world

ORIGIN = (0.0, 0.01, -1600.01)
DIM = (1500.0, 1500.0, 1500.0)


def _centres(res, origin=ORIGIN, dim=DIM):
    """Voxel centres (res, res, res, 3) [z, y, x] in float64."""
    vs = np.array(dim) / res
    start = np.array(origin) - np.array(dim) + vs
    i = np.arange(res)
    X = start[0] + i * 2 * vs[0]
    Y = start[1] + i * 2 * vs[1]
    Z = start[2] + i * 2 * vs[2]
    zz, yy, xx = np.meshgrid(Z, Y, X, indexing="ij")
    return np.stack([xx, yy, zz], -1)


def test_update_plane_closed_form():
    W, H, res, Z = 160, 120, 64, 1600.0
    K = normalized_k(W, H).astype(np.float64)
    g = Grid(res, W, H, ORIGIN, DIM)
    g.update(np.full((H, W), Z, np.float32), np.eye(4), K)
    p = _centres(res)
    D = -p[..., 2]
    u = (p[..., 0] / p[..., 2] * K[0] + K[2]) * W
    v = (p[..., 1] / p[..., 2] * K[4] + K[5]) * H
    sdf = (Z - D) / (DIM[0] * 0.03)
    inside = (np.floor(u) >= 1) & (np.floor(u) < W - 2) & (np.floor(v) >= 1) & (np.floor(v) < H - 2)
    expect_upd = inside & (sdf >= -1.25)
    # leave out voxels within rounding of a decision boundary
    edge = lambda x, b: np.abs(x - b) < 1e-3
    sure = ~(edge(u, 1) | edge(u, W - 2) | edge(v, 1) | edge(v, H - 2) | (np.abs(sdf + 1.25) < 1e-4))
    got = g.data
    assert np.array_equal(got[..., 1][sure] > 0, expect_upd[sure])
    m = sure & expect_upd
    assert m.sum() > 1000
    np.testing.assert_allclose(got[..., 0][m], np.clip(sdf, -1, 1)[m], atol=2e-5)
    assert np.all(got[..., 1][m] == 1.0)
    m0 = sure & ~expect_upd
    assert np.all(got[..., 0][m0] == 1.0) and np.all(got[..., 1][m0] == 0.0)


def test_update_weight_saturates():
    W, H, res = 96, 72, 32
    K = normalized_k(W, H)
    g = Grid(res, W, H, ORIGIN, DIM)
    g.update(np.full((H, W), 1500.0, np.float32), np.eye(4), K)
    once = g.data.copy()
    for _ in range(19):
        g.update(np.full((H, W), 1500.0, np.float32), np.eye(4), K)
    upd = once[..., 1] > 0
    assert np.all(g.data[..., 1][upd] == 16.0)
    np.testing.assert_allclose(g.data[..., 0][upd], once[..., 0][upd], atol=1e-6)
    assert np.all(g.data[..., 1][~upd] == 0.0)


def test_update_rejects_invalid_and_discontinuous_depth():
    """A depth of zero (no measurement) updates nothing; a 100 mm step between
    the four taps (beyond dim / 200) is not fused across."""
    W, H, res = 96, 72, 32
    K = normalized_k(W, H)
    g = Grid(res, W, H, ORIGIN, DIM)
    g.update(np.zeros((H, W), np.float32), np.eye(4), K)
    assert np.all(g.data[..., 1] == 0) and np.all(g.data[..., 0] == 1)
    d = np.full((H, W), 1500.0, np.float32)
    d[:, W // 2:] = 1600.0
    g.update(d, np.eye(4), K)
    p = _centres(res)
    u = (p[..., 0] / p[..., 2] * K[0] + K[2]) * W
    near_step = np.abs(u - (W // 2 - 0.5)) < 0.5  # taps straddle the step
    assert np.all(g.data[..., 1][near_step] == 0)
    assert g.data[..., 1].sum() > 100


def test_raycast_plane():
    W, H, res, Z = 160, 120, 128, 1600.0
    K = normalized_k(W, H)
    g = Grid(res, W, H, ORIGIN, DIM)
    g.update(np.full((H, W), Z, np.float32), np.eye(4), K)
    xyz, nrm, shade = g.raycast(np.eye(4), K, True)
    v = np.zeros((H, W), bool)
    v[4:-4, 4:-4] = True
    assert np.all(xyz[..., 3][v] == 1.0)
    np.testing.assert_allclose(xyz[..., 2][v], -Z, atol=0.02)
    # the point lies on the pixel's ray
    ray = np.stack(np.meshgrid(np.arange(W) / W, np.arange(H) / H), -1)
    np.testing.assert_allclose((xyz[..., 0] / -xyz[..., 2])[v], (-(ray[..., 0] - K[2]) / K[0])[v], atol=1e-5)
    np.testing.assert_allclose((xyz[..., 1] / -xyz[..., 2])[v], (-(ray[..., 1] - K[5]) / K[4])[v], atol=1e-5)
    n = np.zeros((H, W), bool)
    n[5:-5, 5:-5] = True
    np.testing.assert_allclose(nrm[..., 2][n], 1.0, atol=1e-4)
    np.testing.assert_allclose(nrm[..., :2][n], 0.0, atol=1e-2)
    assert np.array_equal(shade[..., 0], nrm[..., 2]) and np.all(shade[..., 3] == 1)
    # the border band the shading leaves out
    assert np.all(nrm[:3] == 0) and np.all(nrm[:, -3:] == 0)


def test_raycast_outside_window_and_missing_rays():
    """Beyond the depth window the point is kept with w = 0; rays that miss
    the cube give zeros."""
    W, H, res = 96, 72, 64
    K = normalized_k(W, H)
    g = Grid(res, W, H, ORIGIN, DIM, min_dist=300.0, max_dist=1000.0)
    g.update(np.full((H, W), 1600.0, np.float32), np.eye(4), K)
    xyz, _, _ = g.raycast(np.eye(4), K, False)
    c = xyz[H // 2, W // 2]
    assert c[3] == 0 and abs(c[2] + 1600) < 0.1
    away = pose(0, 180, (0, 0, 2000))  # behind the cube, looking away
    xyz, _, _ = g.raycast(away, K, True)
    assert np.all(xyz == 0)


@pytest.mark.parametrize("res,p90_surf,p90_depth", [(64, 1.5, 2.0), (128, 0.6, 1.0)])
def test_raycast_scene_from_fused_poses(res, p90_surf, p90_depth):
    W, H = 160, 120
    K = normalized_k(W, H)
    g = Grid(res, W, H, ORIGIN, DIM)
    poses = [pose(), pose(2, -4, (80, -20, 30)), pose(-3, 3, (-60, 40, -50))]
    for T in poses:
        g.update(render_depth(T, K, W, H).astype(np.float32), T, K)
    for T in poses + [pose(1, 1, (20, 10, 0))]:
        xyz, _, _ = g.raycast(T, K, True)
        v = xyz[..., 3] > 0
        v[:4] = v[-4:] = False
        v[:, :4] = v[:, -4:] = False
        assert v.mean() > 0.6
        r = surface_distance(to_world(T, xyz))[v]
        assert np.median(r) < 0.05 and np.percentile(r, 90) < p90_surf
        dz = np.abs(-xyz[..., 2].astype(np.float64) - render_depth(T, K, W, H))[v]
        assert np.median(dz) < 0.05 and np.percentile(dz, 90) < p90_depth


def test_raycast_refinement_moves_toward_the_surface():
    W, H, res = 160, 120, 64
    K = normalized_k(W, H)
    g = Grid(res, W, H, ORIGIN, DIM)
    T = pose(2, -4, (80, -20, 30))
    g.update(render_depth(T, K, W, H).astype(np.float32), T, K)
    errs = []
    for refine in (False, True):
        xyz, _, _ = g.raycast(T, K, refine)
        v = xyz[..., 3] > 0
        errs.append(np.median(surface_distance(to_world(T, xyz))[v]))
    assert errs[1] < 0.5 * errs[0]


@pytest.mark.parametrize("refine", [True, False])
def test_raycast_linear_field(refine):
    """Distance field n.p - c (mm / 100, unclipped): the zero crossing along
    each ray is the plane n.p = c, and the normals are n in camera
    coordinates."""
    W, H, res = 128, 96, 64
    K = normalized_k(W, H)
    g = Grid(res, W, H, ORIGIN, DIM)
    n = np.array([0.3, -0.4, 1.0])
    n /= np.linalg.norm(n)
    c = n @ np.array([0.0, 0.0, -1700.0])
    g.data[..., 0] = (_centres(res) @ n - c) / 100.0
    g.data[..., 1] = 1.0
    T = pose(6, -5, (30, -20, 40))
    xyz, nrm, _ = g.raycast(T, K, refine)
    v = xyz[..., 3] > 0
    v[:4] = v[-4:] = False
    v[:, :4] = v[:, -4:] = False
    assert v.mean() > 0.8
    pw = to_world(T, xyz)
    np.testing.assert_allclose((pw @ n - c)[v], 0.0, atol=0.05)
    ncam = n @ T[:3, :3].astype(np.float64)  # world normal in camera coordinates
    m = nrm[..., 2] != 0
    assert m[6:-6, 6:-6].mean() > 0.8
    np.testing.assert_allclose(nrm[..., :3][m], np.broadcast_to(ncam, nrm[..., :3][m].shape), atol=2e-3)


def test_resample_linear_field():
    res = 32
    g = Grid(res, 64, 48, ORIGIN, DIM)
    p = _centres(res)
    a = np.array([0.3, -0.2, 0.1]) / 100.0
    f = lambda q: (q - np.array(ORIGIN)) @ a
    g.data[..., 0] = f(p)
    g.data[..., 1] = 2.0
    T = pose(3, -5, (40, -25, 60))
    dst = Grid(res, 64, 48, ORIGIN, DIM)
    dst.resample_from(g, T)
    q = p @ T[:3, :3].T.astype(np.float64) + T[:3, 3]
    vs = np.array(DIM) / res
    iv = (q - (np.array(ORIGIN) - np.array(DIM) + vs)) / (2 * vs)
    inr = np.all((iv >= 0) & (iv <= res - 2), -1)
    sure = np.all((np.abs(iv) > 1e-3) & (np.abs(iv - (res - 2)) > 1e-3), -1)
    m = inr & sure
    assert m.sum() > 0.5 * res ** 3
    np.testing.assert_allclose(dst.data[..., 0][m], f(q)[m], atol=1e-5)
    np.testing.assert_allclose(dst.data[..., 1][m], 2.0, atol=1e-5)
    out = ~inr & sure
    assert np.all(dst.data[..., 0][out] == 1) and np.all(dst.data[..., 1][out] == 0)


def test_dist_from_rgbd_closed_form():
    W, H = 40, 30
    g = Grid(8, W, H, ORIGIN, DIM, min_dist=300.0, max_dist=7000.0)
    rng = np.random.default_rng(3)
    rgbd = rng.random((H, W, 4), dtype=np.float32)
    rgbd[0, :5, 3] = [0.0, 1e-4, 1.0, 1 - 1e-4, 0.5]
    d = g.dist_from_rgbd(rgbd)
    wv = rgbd[::-1, :, 3]
    ok = (wv > np.float32(1e-4)) & (wv < np.float32(1 - 1e-4))
    exp = np.where(ok, wv * np.float32(6700.0) + np.float32(300.0), 0).astype(np.float32)
    assert np.array_equal(d, exp)
    assert np.all(d[-1, :4] == 0) and d[-1, 4] == np.float32(0.5) * np.float32(6700) + np.float32(300)


def test_tsdf_create_rejects_bad_params():
    """The C ABI's argument checks (no device work)."""
    from recon_amd import lib
    from recon_amd.tsdf import default_params
    p = default_params(resolution=100)  # not a multiple of 8
    h = ctypes.c_void_p()
    assert lib().rt_tsdf_create(0, ctypes.byref(p), ctypes.byref(h)) == -1
    p = default_params(dim=(0, 1, 1))
    assert lib().rt_tsdf_create(0, ctypes.byref(p), ctypes.byref(h)) == -1
    q = default_params()
    assert q.resolution == 256 and q.width == 320 and q.height == 240
    assert abs(q.origin[2] + 1600.01) < 1e-3 and q.dim[0] == 1500.0


def test_oriented_cloud_keep_rule():
    """The filtered reference's ICP cloud (genOrientedPoints6Kernel): a point
    is kept where z < 0 and |n|^2 > 1e-7, and its record is the point and the
    normal; everything else is zero."""
    import oracle
    W, H = 16, 8
    rng = np.random.default_rng(5)
    xyz = rng.normal(0, 100, (H, W, 4)).astype(np.float32)
    nrm = rng.normal(0, 1, (H, W, 4)).astype(np.float32)
    nrm[0, :4, :3] = [[0, 0, 0], [3e-4, 0, 0], [2e-4, 2e-4, 2e-4], [0, 0, -1]]
    xyz[0, :4, 2] = -5.0
    pts = np.zeros(W * H * 6, np.float32)
    mask = np.zeros(W * H, np.int32)
    oracle.call("or_oriented_cloud_f", xyz, nrm, W, H, pts, mask)
    n2 = (nrm[..., 0] * nrm[..., 0] + nrm[..., 1] * nrm[..., 1] + nrm[..., 2] * nrm[..., 2]).ravel()
    keep = (xyz[..., 2].ravel() < 0) & (n2 > np.float32(1e-7))
    assert np.array_equal(mask.astype(bool), keep)
    assert list(keep[:4]) == [False, False, True, True]  # 9e-8 < 1e-7 < 1.2e-7
    p = pts.reshape(-1, 6)
    assert np.array_equal(p[keep, :3], xyz.reshape(-1, 4)[keep, :3])
    assert np.array_equal(p[keep, 3:], nrm.reshape(-1, 4)[keep, :3])
    assert np.all(p[~keep] == 0)
