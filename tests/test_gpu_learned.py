"""What the trained pose network gains: on unseen large-motion pairs
(rotation <= 10 deg, translation <= 150 mm, 640x480) the batched tracker
started from the network's proposals converges far more often than the same
tracker started from identity. A pair has converged when it ends within
0.1 deg / 1 mm of the pose Gauss-Newton reaches from the true motion (the
anchor). Uses the committed weights recon_amd/data/pose_cnn.npz."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_learned_proposals_raise_the_convergence_rate():
    import bench
    from recon_amd import PHOTOMETRIC
    from recon_amd.cnn import PoseNet
    from recon_amd.synth import calib_block
    from recon_amd.tracker import BatchTracker
    W, H, B = 640, 480, 256
    print(f"case learned gain {B} pairs", flush=True)
    g, d, gt = bench.make_inputs(B, W, H, 10.0, 150.0, 77, nscenes=B)
    npx = W * H
    cal = calib_block(W, H)
    pairs = [dict(ref_gray=g.ptr + 2 * i * npx, ref_depth=d.ptr + 2 * i * npx * 2, cur_gray=g.ptr + (2 * i + 1) * npx,
                  cur_depth=d.ptr + (2 * i + 1) * npx * 2, width=W, height=H, calib=cal) for i in range(B)]
    bt = BatchTracker(B, W, H, mode=PHOTOMETRIC, n_select=8192, iters=(1, 3, 10, 5))
    prep = bt.prepare(pairs)
    anchor, _, _, _ = bt.track(prep, init=list(gt))
    ident, _, _, _ = bt.retrack(np.tile(np.eye(4).ravel(), B), B)
    net = PoseNet.load(max_batch=B)
    bt.propose(net)
    learned, _, _, st = bt.retrack(None, B)

    def rate(T):
        r, t = bench.pose_errors(T, anchor)
        return float(np.mean((r <= 0.1) & (t <= 1.0)))

    ri, rl = rate(ident), rate(learned)
    print(f"converged: identity start {ri:.3f}, learned start {rl:.3f}", flush=True)
    assert rl >= 0.7 and rl >= ri + 0.2, (ri, rl)
