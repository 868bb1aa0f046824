"""CPU restatement of the reference's keyframe search and schedule format.

TEST INFRASTRUCTURE ONLY: the checker for the library's rt_grid_object,
rt_keyframes_similar and rt_schedule_*. FP32 steps are numpy float32
operations in the reference's order.
"""
import numpy as np

F = np.float32


def grid_object(nx, ny, nz, size, view_min, z_range, fov_x_deg):
    """KeyFrameModel::generateGridObject (reconstruct/KeyFrameModel.cpp:707-734)."""
    size, view_min, z_range = F(size), F(view_min), F(z_range)
    sx, sy = (F(2) * size) / F(nx), (F(2) * size) / F(ny)
    sz = F(0)
    if nz == 1:
        view_min = view_min + z_range / F(2)
    else:
        sz = z_range / F(nz - 1)
    tf = np.tan(F(fov_x_deg) * F(3.141592653) / F(180), dtype=np.float32)
    out = []
    for k in range(nz):
        scale = F(1)
        if k > 0:
            z0, zk = -view_min, -view_min - sz * F(k)
            scale = (zk * tf) / (z0 * tf)
        for j in range(ny):
            for i in range(nx):
                out.append(((-size + F(i) * sx) * scale, (-size + F(j) * sy) * scale, -view_min - sz * F(k)))
    return np.array(out, np.float32)


def mat4_mul_f32(A, B):
    """matrixMult4x4 for floats (basic_math.cpp:85-112), left-to-right sums."""
    A, B = np.asarray(A, np.float32).reshape(4, 4), np.asarray(B, np.float32).reshape(4, 4)
    C = np.zeros((4, 4), np.float32)
    for i in range(4):
        for j in range(4):
            C[i, j] = A[i, 0] * B[0, j] + A[i, 1] * B[1, j] + A[i, 2] * B[2, j] + A[i, 3] * B[3, j]
    return C


def invert_rt4_f32(M):
    """invertRT4 for floats (basic_math.cpp:355-375)."""
    M = np.asarray(M, np.float32).ravel()
    Mi = np.zeros(16, np.float32)
    Mi[0], Mi[1], Mi[2] = M[0], M[4], M[8]
    Mi[3] = -(M[0] * M[3] + M[4] * M[7] + M[8] * M[11])
    Mi[4], Mi[5], Mi[6] = M[1], M[5], M[9]
    Mi[7] = -(M[1] * M[3] + M[5] * M[7] + M[9] * M[11])
    Mi[8], Mi[9], Mi[10] = M[2], M[6], M[10]
    Mi[11] = -(M[2] * M[3] + M[6] * M[7] + M[10] * M[11])
    Mi[15] = F(1)
    return Mi.reshape(4, 4)


def pose_distance(dT):
    """poseDistance + rot2euler + rad2deg (basic_math.cpp:35-37, 1059-1094), FP64."""
    d = np.asarray(dT, np.float32).astype(np.float64).ravel()
    dist = float(np.float32(np.sqrt(d[3] ** 2 + d[7] ** 2 + d[11] ** 2 + float(np.float32(1e-12)))))
    t0 = np.arctan2(d[6], d[10])
    t1 = np.arctan2(-d[2], np.sqrt(d[0] * d[0] + d[1] * d[1]))
    s1, c1 = np.sin(t0), np.cos(t0)
    t2 = np.arctan2(s1 * d[8] - c1 * d[4], c1 * d[5] - s1 * d[9])
    a = max(abs(t0), abs(t1))
    a = max(abs(t2), a)
    return dist, float(np.float32(180.0 * float(np.float32(a)) / np.pi))


def projection_distance(relT, grid, K):
    """KeyFrameModel::projectionDistance (KeyFrameModel.cpp:1423-1451), float."""
    R = np.asarray(relT, np.float32).ravel()
    K = np.asarray(K, np.float32).ravel()
    err = F(0)
    for p in np.asarray(grid, np.float32):
        x, y, z = p
        q = [R[0] * x + R[1] * y + R[2] * z + R[3], R[4] * x + R[5] * y + R[6] * z + R[7],
             R[8] * x + R[9] * y + R[10] * z + R[11]]
        a = (x / z, y / z, F(1))
        b = (q[0] / q[2], q[1] / q[2], F(1))
        r1 = (K[0] * a[0] + K[1] * a[1] + K[2] * a[2], K[3] * a[0] + K[4] * a[1] + K[5] * a[2])
        r2 = (K[0] * b[0] + K[1] * b[1] + K[2] * b[2], K[3] * b[0] + K[4] * b[1] + K[5] * b[2])
        dx, dy = r1[0] - r2[0], r1[1] - r2[1]
        err = err + (dx * dx + dy * dy)
    return err / F(len(grid))


def similar_keyframes(T, keys, max_dist, max_angle, K, grid):
    """enumerateNearPoses + findSimilarKeyFrames + bubbleSort
    (KeyFrameModel.cpp:1476-1490, 1530-1600): (indices, weights)."""
    iT = invert_rt4_f32(T)
    close, w = [], []
    wmin, wmax = F(np.finfo(np.float32).max), F(np.finfo(np.float32).tiny)
    for k, P in enumerate(np.asarray(keys, np.float32).reshape(-1, 4, 4)):
        rel = mat4_mul_f32(iT, P)
        d, a = pose_distance(rel)
        if not (np.float32(d) < F(max_dist) and np.float32(a) < F(max_angle)):
            continue
        v = projection_distance(rel, grid, K)
        wmin, wmax = min(wmin, v), max(wmax, v)
        close.append(k)
        w.append(v)
    # bubble sort of the indices by weight, ascending (equal ones keep their order)
    idx = list(range(len(close)))
    for i in range(len(idx) - 1, 0, -1):
        for j in range(1, i + 1):
            if w[idx[j - 1]] > w[idx[j]]:
                idx[j - 1], idx[j] = idx[j], idx[j - 1]
    with np.errstate(invalid="ignore", divide="ignore"):
        weights = [F(1) - (w[i] - wmin) / (wmax - wmin) for i in idx]
    return np.array([close[i] for i in idx], np.int32), np.array(weights, np.float32)


def read_schedule(text):
    """TrackingParams::init (calib/TrackingParams.cpp:50-94): (frame, mode,
    lambda, error) per line; modes 0 geometric, 1 photometric, 2 bi-objective."""
    out = []
    for line in text.splitlines():
        f = line.split()
        if len(f) < 2:
            continue
        name = f[1]
        if name == "geometric":
            m, lam = 0, 1.0
        elif name == "photometric":
            m, lam = 1, 0.0
        elif name.startswith("biobjective"):
            m, lam = 2, float(np.float32(float(name[len("biobjective-lambda"):])))
        else:
            m, lam = -1, 0.0
        out.append((int(f[0]), m, lam, float(np.float32(float(f[2]))) if len(f) > 2 else 0.0))
    return out
