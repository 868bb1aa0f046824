"""The single-pair tracker replays its per-frame Gauss-Newton launches from a
captured graph; the same frames tracked with eager launches (RT_EAGER) give
the same poses bit for bit, in every mode, across keyframe switches, eager
work between the replays, and a change of lambda (which recaptures)."""
import os

import numpy as np
import pytest

from recon_amd import BIOBJECTIVE, GEOMETRIC, PHOTOMETRIC, synth
from recon_amd.tracker import PhotoTracker

pytestmark = pytest.mark.gpu

W, H, N = 320, 240, 12


@pytest.fixture(scope="module")
def frames():
    return synth.make_sequence(3, N, W, H, 1.0, 15.0)


def _run(frames, mode, eager):
    if eager:
        os.environ["RT_EAGER"] = "1"
    else:
        os.environ.pop("RT_EAGER", None)
    try:
        g, d, _ = frames
        tr = PhotoTracker(W, H, synth.calib_block_np(W, H), mode=mode, lam=0.5)
        out = []
        for k in range(1, N):
            if (k - 1) % 4 == 0:
                tr.set_reference_points(g[k - 1], d[k - 1], mode=mode, lam=0.5, force=True)
            if k == 7:
                tr.set_lambda(0.25)
            tr.set_current_image(g[k], d[k])
            assert tr.track()
            out.append(tr.get_global_pose().copy())
            out.append(tr.get_relative_pose().astype(np.float64))
        tr.close()
        return np.array(out)
    finally:
        os.environ.pop("RT_EAGER", None)


@pytest.mark.parametrize("mode", [PHOTOMETRIC, GEOMETRIC, BIOBJECTIVE])
def test_graph_replay_equals_eager_launches(frames, mode):
    print(f"case graph replay mode {mode}", flush=True)
    a = _run(frames, mode, eager=False)
    b = _run(frames, mode, eager=True)
    assert np.array_equal(a, b)
