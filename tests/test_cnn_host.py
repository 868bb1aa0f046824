"""Host parts of the pose network's C ABI (no GPU): pose <-> target
conversions, configuration checks and the parameter layout."""
import ctypes

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from recon_amd import lib
from recon_amd.cnn import default_config, init_params, layout, pose_to_target, target_to_pose
from recon_amd.synth import sample_motion


def test_default_network_and_layout():
    c = default_config()
    assert (c.in_width, c.in_height) == (80, 60)
    assert list(c.channels) == [4, 32, 64, 128, 128, 256, 256]
    layers, fc_w, fc_b, total = layout(c)
    assert total == 1127910
    assert layers[1] == (4 * 32 * 9 + 32, 4 * 32 * 9 + 32 + 64 * 9 * 32, 64, 32)
    p = init_params(c, 0)
    assert p.size == total and np.all(p[layers[0][1]:layers[0][1] + 32] == 0)


def test_targets_round_trip_and_match_rotation_vectors():
    c = default_config()
    T = np.array([np.linalg.inv(sample_motion(s, 10.0, 150.0)) for s in range(300)])
    y = pose_to_target(c, T)
    rv = Rotation.from_matrix(T[:, :3, :3]).as_rotvec()
    assert np.allclose(y[:, :3], rv / np.radians(10.0), atol=1e-6)
    assert np.allclose(y[:, 3:], T[:, :3, 3] / 150.0, atol=1e-6)
    assert np.abs(y).max() <= 1.0 + 1e-6
    T2 = target_to_pose(c, y)
    assert np.allclose(T2, T, atol=1e-6)


def test_near_pi_and_zero_rotations():
    c = default_config(rot_scale_deg=180.0)
    for rv in ([0.0, 0.0, 0.0], [1e-12, 0, 0], [np.pi - 1e-8, 0, 0], [0, 0.3, -0.2]):
        T = np.eye(4)
        T[:3, :3] = Rotation.from_rotvec(rv).as_matrix()
        y = pose_to_target(c, T[None])
        assert np.allclose(np.abs(y[0, :3] * np.pi), np.abs(rv), atol=1e-6)


@pytest.mark.parametrize("kw", [dict(channels=[3, 32, 64, 128, 128, 256, 256]), dict(channels=[4, 24, 64, 128, 128, 256, 256]),
                                dict(rot_scale_deg=0.0)])
def test_create_rejects_bad_configs_before_touching_the_device(kw):
    c = default_config(**kw)
    h = ctypes.c_void_p()
    assert lib().rt_cnn_create(0, ctypes.byref(c), 16, ctypes.byref(h)) == -1
    assert not h.value
