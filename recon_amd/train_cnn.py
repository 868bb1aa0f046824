"""Train the pose-proposal network on the GPU.

Data: seeded synthetic frame pairs at 640x480 (a new random scene per pair,
rotation up to --theta degrees, translation up to --tmax mm), rendered on the
GPU. Labels are the poses the plain Gauss-Newton tracker converges to when
started at the true motion (same kernels, batched), kept where it converged
(status 0 and within 1 deg / 20 mm of the truth). Inputs are the network's
80x60 4-channel tensors built on the GPU; the training steps (forward,
backward, Adam) are the library's own kernels.

    python -m recon_amd.train_cnn --pairs 200000 --epochs 12 --out recon_amd/data/pose_cnn.npz
"""
import argparse
import ctypes
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from recon_amd import PHOTOMETRIC, synth_lib  # noqa: E402
from recon_amd.cnn import PoseNet, default_config, init_params, pair_array  # noqa: E402
from recon_amd.hip import DeviceBuffer, hip  # noqa: E402
from recon_amd.synth import calib_block, pinhole, pose_error, sample_motion  # noqa: E402
from recon_amd.tracker import BatchTracker  # noqa: E402


def log(*a):
    print(*a, flush=True)


def motion_seed(s):
    return (int(s) * 0x9E3779B97F4A7C15 + 0x632BE59BD9B4E019) % (1 << 64)


class PairRenderer:
    """Renders chunks of seeded frame pairs into reusable device buffers."""

    def __init__(self, chunk, W=640, H=480):
        self.S = synth_lib()
        self.W, self.H, self.chunk = W, H, chunk
        self.g = DeviceBuffer(2 * chunk * W * H)
        self.d = DeviceBuffer(2 * chunk * W * H * 2)
        self.intr = pinhole(W, H)

    def render(self, seeds, theta, tmax):
        """Pairs for the given seeds (one scene and one motion per seed);
        returns the pair dicts and the true Trel of each."""
        W, H = self.W, self.H
        npx = W * H
        dp = ctypes.POINTER(ctypes.c_double)
        fx, fy, cx, cy = self.intr
        gts, pairs = [], []
        cal = calib_block(W, H)
        for i, s in enumerate(seeds):
            sc = self.S.rs_scene_create(ctypes.c_uint64(int(s)))
            # the scene and the motion draw from one generator family: give the
            # motion its own seed so that it is independent of the scene
            M = sample_motion(motion_seed(int(s)), theta, tmax)
            for f, T in ((0, np.eye(4)), (1, M)):
                k = 2 * i + f
                Tf = np.ascontiguousarray(T, np.float64).ravel()
                rc = self.S.rs_render_device(sc, Tf.ctypes.data_as(dp), W, H, fx, fy, cx, cy,
                                             ctypes.c_uint64(int(s) * 2 + f), 0.01, self.g.ptr + k * npx,
                                             self.d.ptr + k * npx * 2, None)
                if rc != 0:
                    raise RuntimeError("render failed")
            self.S.rs_scene_destroy(sc)
            gts.append(np.linalg.inv(M))
            pairs.append(dict(ref_gray=self.g.ptr + 2 * i * npx, ref_depth=self.d.ptr + 2 * i * npx * 2,
                              cur_gray=self.g.ptr + (2 * i + 1) * npx, cur_depth=self.d.ptr + (2 * i + 1) * npx * 2,
                              width=W, height=H, calib=cal))
        hip().hipDeviceSynchronize()
        return pairs, np.array(gts)


def make_dataset(net, renderer, tracker, n, seed0, theta, tmax, X, Y, Yinv, Wt, offset=0):
    """Fills X (inputs), Y (targets), Yinv (targets of the inverse poses) and
    Wt (sample weights) for n pairs from seed0 on; returns the true and the
    label poses."""
    chunk = renderer.chunk
    elems = net.input_elems
    gts, labels = [], []
    for c0 in range(0, n, chunk):
        m = min(chunk, n - c0)
        seeds = seed0 + c0 + np.arange(m)
        pairs, gt = renderer.render(seeds, theta, tmax)
        net.make_inputs(pair_array(pairs), m, X.ptr + 2 * elems * (offset + c0))
        Trel, _, _, st = tracker.track(pairs, init=list(gt))
        err = np.array([pose_error(a, b) for a, b in zip(Trel, gt)])
        w = ((st == 0) & (err[:, 0] < 1.0) & (err[:, 1] < 20.0)).astype(np.float32)
        y = net.pose_to_target(Trel)
        yi = net.pose_to_target(np.linalg.inv(Trel))
        y[w == 0] = 0.0
        yi[w == 0] = 0.0
        hip().hipMemcpy(Y.ptr + 24 * (offset + c0), y.ctypes.data, y.nbytes, 1)
        hip().hipMemcpy(Yinv.ptr + 24 * (offset + c0), yi.ctypes.data, yi.nbytes, 1)
        hip().hipMemcpy(Wt.ptr + 4 * (offset + c0), w.ctypes.data, w.nbytes, 1)
        gts.append(gt); labels.append(Trel)
        log(f"data: {c0 + m}/{n} pairs, {w.mean() * 100:.1f}% labelled")
    return np.concatenate(gts), np.concatenate(labels)


def evaluate(net, X, n, offset, gt, batch=4096):
    out = DeviceBuffer(24 * batch)
    props = []
    for c0 in range(0, n, batch):
        m = min(batch, n - c0)
        net.forward(m, X.ptr + 2 * net.input_elems * (offset + c0), out.ptr)
        hip().hipDeviceSynchronize()
        props.append(net.target_to_pose(out.to_array(np.float32, 6 * m)))
    P = np.concatenate(props)
    e = np.array([pose_error(a, b) for a, b in zip(P, gt)])
    return P, e


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=200000)
    ap.add_argument("--val", type=int, default=4096)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--theta", type=float, default=10.0)
    ap.add_argument("--tmax", type=float, default=150.0)
    ap.add_argument("--chunk", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=1_000_000)
    ap.add_argument("--out", default=os.path.join(ROOT, "recon_amd", "data", "pose_cnn.npz"))
    ap.add_argument("--time-limit", type=float, default=0.0, help="stop training after this many seconds (0: none)")
    ap.add_argument("--no-augment", action="store_true", help="train without the mirror / swap symmetries")
    args = ap.parse_args()

    t0 = time.time()
    cfg = default_config()
    B = args.batch
    net = PoseNet(max(B, args.chunk), config=cfg)
    net.set_params(init_params(cfg, 0))
    ntot = args.pairs + args.val
    X = DeviceBuffer(2 * net.input_elems * ntot)
    Y = DeviceBuffer(24 * ntot)
    Yinv = DeviceBuffer(24 * ntot)
    Wt = DeviceBuffer(4 * ntot)
    Xa = DeviceBuffer(2 * net.input_elems * B)
    Ya = DeviceBuffer(24 * B)
    renderer = PairRenderer(args.chunk)
    tracker = BatchTracker(args.chunk, 640, 480, mode=PHOTOMETRIC, n_select=8192, iters=(1, 3, 10, 5))
    make_dataset(net, renderer, tracker, args.pairs, args.seed, args.theta, args.tmax, X, Y, Yinv, Wt, 0)
    vgt, vlab = make_dataset(net, renderer, tracker, args.val, args.seed + 50_000_000, args.theta, args.tmax, X, Y, Yinv,
                             Wt, args.pairs)
    log(f"dataset ready in {time.time() - t0:.0f}s")

    steps_per_epoch = args.pairs // B
    total = steps_per_epoch * args.epochs
    rng = np.random.default_rng(args.seed)
    elems = net.input_elems
    step = 0
    t1 = time.time()
    for ep in range(args.epochs):
        off = int(rng.integers(0, B))
        losses = []
        for s in range(steps_per_epoch):
            start = (off + s * B) % (args.pairs - B + 1)
            lr = args.lr * 0.5 * (1.0 + np.cos(np.pi * step / total)) + 1e-6
            want = (s % 100 == 0)
            xs, ys = X.ptr + 2 * elems * start, Y.ptr + 24 * start
            if not args.no_augment:
                net.augment(B, xs, ys, Yinv.ptr + 24 * start, rng.integers(0, 8, B), Xa.ptr, Ya.ptr)
                xs, ys = Xa.ptr, Ya.ptr
            l = net.train_step(B, xs, ys, Wt.ptr + 4 * start, lr, want_loss=want)
            if want:
                losses.append(l)
            step += 1
        _, e = evaluate(net, X, args.val, args.pairs, vgt)
        log(f"epoch {ep + 1}/{args.epochs}: train loss {np.mean(losses):.4f}, val proposal error median "
            f"{np.median(e[:, 0]):.3f} deg {np.median(e[:, 1]):.2f} mm, within 3deg/50mm "
            f"{np.mean((e[:, 0] < 3) & (e[:, 1] < 50)) * 100:.1f}% ({time.time() - t1:.0f}s)")
        if args.time_limit and time.time() - t0 > args.time_limit:
            log("time limit reached")
            break
    net.save(args.out)
    log(f"saved {args.out} ({time.time() - t0:.0f}s total)")


if __name__ == "__main__":
    main()
