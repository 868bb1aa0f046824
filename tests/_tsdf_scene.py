"""An analytic scene for the TSDF tests: a sphere in front of a tilted wall, seen by
a pinhole camera in the reference's conventions (camera looks down -z, depth
in mm, normalised intrinsics K = [fx/w, 0, cx/w; 0, fy/h, cy/h; 0, 0, 1] with
fx < 0, pixel (x, y) at normalised position (x/w, y/h))."""
import numpy as np

SPHERE_C = np.array([120.0, -60.0, -1500.0])
SPHERE_R = 350.0
WALL_N = np.array([0.25, 0.1, 1.0]) / np.linalg.norm([0.25, 0.1, 1.0])  # unit normal
WALL_P = np.array([0.0, 0.0, -2400.0])  # a point of the wall


def normalized_k(w, h):
    f = 525.0 * w / 640.0
    return np.array([-f / w, 0, 0.5, 0, f / h, 0.5, 0, 0, 1], np.float32)


def pose(rx_deg=0.0, ry_deg=0.0, t=(0.0, 0.0, 0.0)):
    """Camera -> world: rotation about x then y, then translation (mm)."""
    a, b = np.radians(rx_deg), np.radians(ry_deg)
    Rx = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
    Ry = np.array([[np.cos(b), 0, np.sin(b)], [0, 1, 0], [-np.sin(b), 0, np.cos(b)]])
    T = np.eye(4)
    T[:3, :3] = Ry @ Rx
    T[:3, 3] = t
    return T.astype(np.float32)


def camera_rays(w, h, K):
    """Unnormalised camera-frame ray per pixel with z = -1, (h, w, 3) float64."""
    K = np.asarray(K, np.float64)
    xs = np.arange(w) / w
    ys = np.arange(h) / h
    X, Y = np.meshgrid(xs, ys)
    return np.stack([-(X - K[2]) / K[0], -(Y - K[5]) / K[4], -np.ones_like(X)], -1)


def render_depth(T, K, w, h):
    """z-depth (mm, float64) of the scene from camera pose T; 0 where nothing is hit."""
    T = np.asarray(T, np.float64)
    d = camera_rays(w, h, K) @ T[:3, :3].T  # world direction; camera depth along the ray = t
    o = T[:3, 3]
    oc = o - SPHERE_C
    b = d @ oc
    a = (d * d).sum(-1)
    c = oc @ oc - SPHERE_R ** 2
    disc = b * b - a * c
    ts = np.where(disc > 0, (-b - np.sqrt(np.maximum(disc, 0))) / a, np.inf)
    ts = np.where(ts > 0, ts, np.inf)
    with np.errstate(divide="ignore", invalid="ignore"):
        tw = ((WALL_P - o) @ WALL_N) / (d @ WALL_N)
    tw = np.where(tw > 0, tw, np.inf)
    t = np.minimum(ts, tw)
    return np.where(np.isfinite(t), t, 0.0)


def surface_distance(p):
    """Distance of world points (..., 3) to the nearest scene surface."""
    ds = np.abs(np.linalg.norm(p - SPHERE_C, axis=-1) - SPHERE_R)
    dw = np.abs((p - WALL_P) @ WALL_N)
    return np.minimum(ds, dw)


def to_world(T, xyz):
    T = np.asarray(T, np.float64)
    return xyz[..., :3].astype(np.float64) @ T[:3, :3].T + T[:3, 3]
