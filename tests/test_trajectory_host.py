"""Trajectory and schedule files and the similar-keyframe search on the CPU:
the library's trajectory file against the reference's own writer byte for
byte, the reference's sample files read back, poseDistance against the
reference's golden vectors, and the grid object and keyframe search against
the oracle."""
import os

import numpy as np
import pytest

from oracle import keyframes as okf
from recon_amd import BIOBJECTIVE, GEOMETRIC, PHOTOMETRIC, trajectory

HERE = os.path.dirname(os.path.abspath(__file__))
G = os.path.join(HERE, "golden")
GOLD = np.load(os.path.join(G, "ref_pose_math.npz"))


@pytest.mark.parametrize("transpose", [0, 1])
@pytest.mark.parametrize("inverse", [0, 1])
def test_trajectory_file_equals_the_reference_writer(tmp_path, transpose, inverse):
    poses = np.load(os.path.join(G, "ref_trajectory_poses.npy"))
    p = tmp_path / "traj.txt"
    trajectory.write_trajectory(p, poses, transpose=transpose, inverse=inverse)
    assert p.read_bytes() == open(os.path.join(G, f"ref_trajectory_t{transpose}_i{inverse}.txt"), "rb").read()
    # read back: the canonised poses (relative to the first) to the file's precision
    back = trajectory.read_trajectory(p, transpose=transpose, inverse=inverse)
    P = poses.reshape(-1, 4, 4).astype(np.float64)
    canon = np.linalg.inv(P[0]) @ P
    assert back.shape == canon.shape
    assert np.allclose(back, canon, rtol=2e-6, atol=2e-6 * np.abs(canon).max())


def test_reference_sample_trajectory_reads():
    back = trajectory.read_trajectory(os.path.join(G, "cameraMatrixEst.txt"))
    assert back.shape == (16, 4, 4) and np.array_equal(back, np.tile(np.eye(4, dtype=np.float32), (16, 1, 1)))


def test_schedule_reads_the_reference_sample():
    path = os.path.join(G, "best-method.tst")
    entries, n = trajectory.read_schedule(path)
    want = okf.read_schedule(open(path).read())
    assert n == len(want) == 118
    got = [(entries[i].frame, entries[i].mode, entries[i].lambda_, entries[i].error) for i in range(n)]
    assert got == want
    assert got[0] == (1, PHOTOMETRIC, 0.0, float(np.float32(246.0214)))
    assert got[2][:3] == (60, BIOBJECTIVE, float(np.float32(0.75)))
    assert any(g[1] == GEOMETRIC and g[2] == 1.0 for g in got)
    # nearest entry, the first of equally near ones (TrackingParams.cpp:96-110)
    assert trajectory.schedule_lookup(entries, n, 45) == (PHOTOMETRIC, 0.0)  # 30 and 60 are equally near
    assert trajectory.schedule_lookup(entries, n, 52) == (BIOBJECTIVE, np.float32(0.75))
    assert trajectory.schedule_lookup(entries, 0, 10) == (GEOMETRIC, 1.0)


def test_pose_distance_equals_the_reference():
    for T, (d, a) in zip(GOLD["pd_T"], GOLD["pd_dist_angle"]):
        assert trajectory.pose_distance(T) == (d, a)
        assert okf.pose_distance(T) == (d, a)


def test_grid_object_matches_oracle_and_geometry():
    g = trajectory.grid_object(30, 30, 5, 1000.0, 2000.0, 3000.0, 62.0)
    o = okf.grid_object(30, 30, 5, 1000.0, 2000.0, 3000.0, 62.0)
    assert g.shape == (4500, 3) and np.array_equal(g, o)
    z = np.unique(g[:, 2])
    assert np.allclose(sorted(z), [-5000, -4250, -3500, -2750, -2000])
    far = g[g[:, 2] == -5000]
    assert np.allclose(far[:, 0].min(), -1000 * 2.5, rtol=1e-6)  # the square grows with depth
    one = trajectory.grid_object(4, 4, 1, 500.0, 1000.0, 1000.0, 50.0)
    assert np.all(one[:, 2] == -1500) and np.array_equal(one, okf.grid_object(4, 4, 1, 500.0, 1000.0, 1000.0, 50.0))


def _pose(rotvec, t):
    from scipy.spatial.transform import Rotation
    T = np.eye(4)
    T[:3, :3] = Rotation.from_rotvec(rotvec).as_matrix()
    T[:3, 3] = t
    return T.astype(np.float32)


def test_similar_keyframes_match_oracle():
    rng = np.random.default_rng(9)
    grid = trajectory.grid_object(6, 6, 3, 1000.0, 2000.0, 3000.0, 60.0)
    K = np.array([-525, 0, 319.5, 0, 525, 239.5, 0, 0, 1], np.float32)
    T = _pose([0.1, -0.2, 0.05], [100, 50, -30])
    keys = [T]  # the pose itself
    for i in range(40):
        keys.append(T @ _pose(rng.normal(0, 0.15, 3), rng.normal(0, 300, 3)))
    keys = np.array(keys)
    idx, w = trajectory.similar_keyframes(T, keys, 500.0, 20.0, K, grid)
    oidx, ow = okf.similar_keyframes(T, keys, 500.0, 20.0, K, grid)
    assert np.array_equal(idx, oidx) and np.array_equal(w, ow)
    assert idx[0] == 0 and w[0] == 1.0 and 1 < len(idx) < 41
    # every one found is within the thresholds, every one left out is not
    for k in range(len(keys)):
        d, a = trajectory.pose_distance(np.linalg.inv(T.astype(np.float64)) @ keys[k])
        assert (k in idx) == (d < 500.0 and a < 20.0) or abs(d - 500) < 1e-3 or abs(a - 20) < 1e-3
