"""Inspection images of the tracker (numpy float32 restatement).

TEST INFRASTRUCTURE ONLY: the checker of the HIP kernels behind
rt_update_error_image / rt_get_debug_image / rt_get_ref_image. Every value is
computed with the reference's FP32 operations in its order (numpy float32
arithmetic is IEEE single, one rounding per operation).
"""
import numpy as np

f32 = np.float32
MIND, MAXD, KL = 62, 63, 43


def ref_depth_image(z, calib):
    """xyz2DepthNormalizedKernel (tracker/kernels/rgb2gray.cu:104-116) over the
    points of convertZMapToXYZKernel (:570-593): a pixel's z if
    -max < z < -min, else 0; then clamp(-z, min, max) mapped to [0, 1]."""
    z = np.asarray(z, f32)
    mn, mx = f32(calib[MIND]), f32(calib[MAXD])
    pz = np.where((z < -mn) & (z > -mx), z, f32(0))
    zval = np.clip(-pz, mn, mx).astype(f32)
    return ((zval - mn) / (mx - mn)).astype(f32)


def photo_error_image(sel, residuals, n_pixels, out=None):
    """xyz2DiffKernel (tracker/kernels/warp.cu:1955-1999): at each selected
    pixel min(25 |r|, 1) of its level-0 residual (r = 1 outside the view);
    the other pixels of `out` (zeros by default) are left as they are."""
    out = np.zeros(n_pixels, f32) if out is None else out
    r = np.asarray(residuals, f32)
    out[np.asarray(sel)] = np.minimum(np.abs(r) * f32(25.0), f32(1.0))
    return out


def icp_error_image(ref_pts, ref_mask, cur_pts, cur_mask, w, h, T, calib, layer):
    """xyz6DiffNormalizedKernel (tracker/kernels/rgb2gray.cu:143-171) with the
    level's a = 2^-l, b = (a - 1) / 2 (pointCloud6DiffCuda, :745-760): per
    keyframe pixel min(|z' - z_cur| / 30, 1) against the current point at the
    nearest pixel its transform projects to with KL; 0 where the keyframe or
    the current pixel is not selected or the projection leaves the view.
    Points are 6-float records (x, y, z, nx, ny, nz)."""
    T = np.asarray(T, f32).ravel()
    K = np.asarray(calib, f32)[KL:KL + 9]
    a = f32(1.0) / f32(1 << layer)
    b = f32(0.5) * (a - f32(1.0))
    P = np.asarray(ref_pts, f32).reshape(-1, 6)
    x, y, z = P[:, 0], P[:, 1], P[:, 2]
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        q = [((T[4 * r] * x + T[4 * r + 1] * y) + T[4 * r + 2] * z) + T[4 * r + 3] for r in range(3)]
        qx, qy = q[0] / q[2], q[1] / q[2]
        u = a * (K[0] * qx + K[2]) + b
        v = a * (K[4] * qy + K[5]) + b
        inview = (u >= 0) & (v >= 0) & (u < f32(w - 1)) & (v < f32(h - 1))
        ui = np.where(inview, u + f32(0.5), 0).astype(np.int32)
        vi = np.where(inview, v + f32(0.5), 0).astype(np.int32)
    off2 = ui + vi * w
    C = np.asarray(cur_pts, f32).reshape(-1, 6)
    cm = np.asarray(cur_mask) > 0
    hit = (np.asarray(ref_mask) != 0) & inview & cm[off2]
    d = np.minimum(np.abs(q[2] - C[off2, 2]) / f32(30.0), f32(1.0))
    return np.where(hit, d, f32(0)).astype(f32)
