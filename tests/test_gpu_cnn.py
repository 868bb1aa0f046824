"""Pose network kernels on the GPU against the CPU oracle (oracle/cnn.hpp with
bfloat16 storage emulated): forward outputs, loss and every parameter
gradient; the Adam update against its closed form; the input builder against
numpy; the batch proposals against forward + host conversion.

Tolerances: the device accumulates in FP32 where the oracle uses FP64, so a
stored bfloat16 value can round the other way where it lies within ~1e-6 of a
rounding boundary; outputs are compared within 2e-3 of their scale, gradients
within 2e-2 relative (Frobenius norm per tensor)."""
import ctypes

import numpy as np
import pytest

from oracle import lib as olib, ptr
from recon_amd.cnn import PoseNet, default_config, init_params, layout, pair_array, target_to_pose
from recon_amd.hip import DeviceBuffer

pytestmark = pytest.mark.gpu

SMALL = dict(channels=[4, 16, 16, 32, 32, 16, 16], in_width=18, in_height=13)


def bf16_bits(x):
    r = np.zeros_like(x, np.float32)
    olib().or_bf16_round(ptr(np.ascontiguousarray(x, np.float32)), x.size, ptr(r))
    return r, (r.view(np.uint32) >> 16).astype(np.uint16)


def oracle_fb(cfg, params, x, y, wn, grad=True):
    L = olib()
    L.or_cnn_forward_backward.restype = ctypes.c_double
    ch = np.array(list(cfg.channels), np.int32)
    B = x.shape[0]
    out = np.zeros(B * 6)
    g = np.zeros(params.size) if grad else None
    loss = L.or_cnn_forward_backward(ptr(ch), cfg.in_height, cfg.in_width, B, ptr(params), ptr(x), ptr(y), ptr(wn), 1,
                                     ptr(out), ptr(g) if grad else None)
    return loss, out.reshape(B, 6), g


def case(cfg, B, seed):
    rng = np.random.default_rng(seed)
    p = init_params(cfg, seed)
    layers, fc_w, fc_b, total = layout(cfg)
    for (w, b, co, ci) in layers:
        p[b:b + co] = rng.normal(0, 0.05, co)
    p[fc_w:total] = rng.normal(0, 0.3, total - fc_w)
    x = rng.normal(0, 0.6, (B, cfg.in_height, cfg.in_width, 4)).astype(np.float32)
    xr, xb = bf16_bits(x)
    y = rng.normal(0, 1, (B, 6)).astype(np.float32)
    wn = np.ones(B, np.float32)
    if B > 2:
        wn[1] = 0.0
    return p, xr, xb, y, wn


@pytest.mark.parametrize("kw,B", [({}, 3), ({}, 37), (SMALL, 5)])
def test_forward_matches_oracle(kw, B):
    print(f"case cnn forward {kw} B={B}", flush=True)
    cfg = default_config(**kw)
    net = PoseNet(max(8, B), config=cfg)
    p, xr, xb, y, wn = case(cfg, B, 1)
    net.set_params(p)
    xd = DeviceBuffer.from_array(xb)
    od = DeviceBuffer(24 * B)
    net.forward(B, xd.ptr, od.ptr)
    from recon_amd.hip import hip
    hip().hipDeviceSynchronize()
    out = od.to_array(np.float32, 6 * B).reshape(B, 6)
    _, ref, _ = oracle_fb(cfg, p, xr, y, wn, grad=False)
    scale = max(1.0, float(np.abs(ref).max()))
    assert np.abs(out - ref).max() <= 2e-3 * scale, (out, ref)


@pytest.mark.parametrize("kw,B", [({}, 3), (SMALL, 5)])
def test_gradients_match_oracle(kw, B):
    print(f"case cnn gradients {kw} B={B}", flush=True)
    cfg = default_config(**kw)
    net = PoseNet(8, config=cfg)
    p, xr, xb, y, wn = case(cfg, B, 2)
    net.set_params(p)
    xd, yd, wd = DeviceBuffer.from_array(xb), DeviceBuffer.from_array(y), DeviceBuffer.from_array(wn)
    loss, out, g = net.loss_grad(B, xd.ptr, yd.ptr, wd.ptr)
    rl, rout, rg = oracle_fb(cfg, p, xr, y, wn)
    assert abs(loss - rl) <= 2e-3 * max(1.0, abs(rl))
    assert np.abs(out - rout).max() <= 2e-3 * max(1.0, float(np.abs(rout).max()))
    layers, fc_w, fc_b, total = layout(cfg)
    spans = [(f"W{l}", w, b) for l, (w, b, co, ci) in enumerate(layers)] + \
            [(f"b{l}", b, b + co) for l, (w, b, co, ci) in enumerate(layers)] + [("Wfc", fc_w, fc_b), ("bfc", fc_b, total)]
    for name, a, b in spans:
        ref = rg[a:b]
        err = np.linalg.norm(g[a:b] - ref) / max(np.linalg.norm(ref), 1e-12)
        assert err < 2e-2, (name, err)


def test_adam_step_matches_closed_form():
    print("case cnn adam", flush=True)
    cfg = default_config(**SMALL)
    net = PoseNet(8, config=cfg)
    B = 4
    p, xr, xb, y, wn = case(cfg, B, 3)
    net.set_params(p)
    xd, yd, wd = DeviceBuffer.from_array(xb), DeviceBuffer.from_array(y), DeviceBuffer.from_array(wn)
    _, _, g1 = net.loss_grad(B, xd.ptr, yd.ptr, wd.ptr)
    lr = 1e-3
    l0 = net.train_step(B, xd.ptr, yd.ptr, wd.ptr, lr)
    p1 = net.get_params()
    m = 0.1 * g1.astype(np.float64); v = 0.001 * g1.astype(np.float64) ** 2
    expect = p - lr * (m / 0.1) / (np.sqrt(v / 0.001) + 1e-8)
    assert np.allclose(p1, expect, rtol=1e-5, atol=1e-6)
    # a few more steps on the same batch lower its loss
    for _ in range(60):
        l = net.train_step(B, xd.ptr, yd.ptr, wd.ptr, 2e-3)
    assert l < 0.6 * l0, (l0, l)


def test_inputs_from_frames():
    print("case cnn inputs", flush=True)
    from recon_amd import synth
    W, H = 160, 120
    g0, d0, g1, d1, _ = synth.make_pair(4, W, H, 5.0, 80.0)
    d1 = d1.copy()
    d1[:16, :16] = 0  # a block with no depth
    bufs = [DeviceBuffer.from_array(a) for a in (g0, d0, g1, d1)]
    net = PoseNet(2)
    x = DeviceBuffer(2 * 80 * 60 * 4)
    arr = pair_array([dict(ref_gray=bufs[0].ptr, ref_depth=bufs[1].ptr, cur_gray=bufs[2].ptr, cur_depth=bufs[3].ptr,
                           width=W, height=H)])
    net.make_inputs(arr, 1, x.ptr)
    got = (x.to_array(np.uint16, 80 * 60 * 4).astype(np.uint32) << 16).view(np.float32).reshape(60, 80, 4)
    ref = np.zeros((60, 80, 4))
    for f, (g, d) in enumerate(((g0, d0), (g1, d1))):
        gb = g.reshape(60, 2, 80, 2).astype(np.float64).mean(axis=(1, 3)) / 255.0 - 0.5
        dd = d.reshape(60, 2, 80, 2).astype(np.float64)
        inv = np.where(dd > 0, 1000.0 / np.maximum(dd, 1), 0.0)
        nv = (dd > 0).sum(axis=(1, 3))
        ib = np.where(nv > 0, inv.sum(axis=(1, 3)) / np.maximum(nv, 1) - 1.0, -1.0)
        ref[..., 2 * f], ref[..., 2 * f + 1] = gb, ib
    # one bfloat16 step of the value
    assert np.all(np.abs(got - ref) <= np.abs(ref) * 2.0 ** -8 + 1e-6)
    assert np.all(got[:8, :8, 3] == -1.0)


def test_batch_proposals_are_the_network_outputs():
    """rt_batch_propose writes Trel = pose(forward(inputs)); with zero
    Gauss-Newton iterations the tracking launch returns it unchanged."""
    print("case cnn propose", flush=True)
    from recon_amd import PHOTOMETRIC, synth
    from recon_amd.tracker import BatchTracker
    W, H, n = 160, 120, 3
    frames = [synth.make_pair(10 + i, W, H, 8.0, 120.0) for i in range(n)]
    bufs = [[DeviceBuffer.from_array(a) for a in f[:4]] for f in frames]
    cal = synth.calib_block_np(W, H)
    pairs = [dict(ref_gray=b[0].ptr, ref_depth=b[1].ptr, cur_gray=b[2].ptr, cur_depth=b[3].ptr, width=W, height=H,
                  calib=cal) for b in bufs]
    cfg = default_config()
    net = PoseNet(8, config=cfg)
    p = init_params(cfg, 5)
    layers, fc_w, fc_b, total = layout(cfg)
    p[fc_w:total] = np.random.default_rng(1).normal(0, 0.5, total - fc_w)
    net.set_params(p)
    x = DeviceBuffer(2 * net.input_elems * n)
    o = DeviceBuffer(24 * n)
    net.make_inputs(pair_array(pairs), n, x.ptr)
    net.forward(n, x.ptr, o.ptr)
    from recon_amd.hip import hip
    hip().hipDeviceSynchronize()
    expect = target_to_pose(cfg, o.to_array(np.float32, 6 * n))
    bt = BatchTracker(n, W, H, mode=PHOTOMETRIC, n_select=2048, iters=(0, 0, 0, 0))
    prep = bt.prepare(pairs)
    bt.prepare_only(prep)
    bt.propose(net)
    Trel, _, _, st = bt.retrack(None, n)
    assert np.all(st == 0)
    assert np.allclose(Trel, expect, atol=2e-6), np.abs(Trel - expect).max()


def test_batch_inputs_built_in_preparation():
    """Once rt_batch_propose has been called on a batch, the preparation
    builds the network's inputs itself (the keyframe's in the gradient-bin
    pass, the current frame's in its pyramid pass) instead of the network's
    input builder; the proposals must not change by a bit. 640x480 pairs
    (80x60 cells of 8x8), with depth-less blocks in either frame."""
    print("case cnn inputs built in the preparation", flush=True)
    from recon_amd import PHOTOMETRIC, synth
    from recon_amd.tracker import BatchTracker
    W, H, n = 640, 480, 3
    frames = [list(synth.make_pair(30 + i, W, H, 8.0, 120.0)[:4]) for i in range(n)]
    frames[0][1] = frames[0][1].copy(); frames[0][1][:24, :40] = 0     # keyframe depth
    frames[1][3] = frames[1][3].copy(); frames[1][3][100:117, 300:333] = 0  # current depth
    bufs = [[DeviceBuffer.from_array(a) for a in f] for f in frames]
    cal = synth.calib_block_np(W, H)
    pairs = [dict(ref_gray=b[0].ptr, ref_depth=b[1].ptr, cur_gray=b[2].ptr, cur_depth=b[3].ptr, width=W, height=H,
                  calib=cal) for b in bufs]
    cfg = default_config()
    assert cfg.in_width * 8 == W and cfg.in_height * 8 == H
    net = PoseNet(8, config=cfg)
    p = init_params(cfg, 6)
    layers, fc_w, fc_b, total = layout(cfg)
    p[fc_w:total] = np.random.default_rng(2).normal(0, 0.5, total - fc_w)
    net.set_params(p)
    x = DeviceBuffer(2 * net.input_elems * n)
    o = DeviceBuffer(24 * n)
    net.make_inputs(pair_array(pairs), n, x.ptr)
    net.forward(n, x.ptr, o.ptr)
    from recon_amd.hip import hip
    hip().hipDeviceSynchronize()
    expect = target_to_pose(cfg, o.to_array(np.float32, 6 * n))
    bt = BatchTracker(n, W, H, mode=PHOTOMETRIC, n_select=2048, iters=(0, 0, 0, 0))
    prep = bt.prepare(pairs)
    got = []
    for step in range(3):  # the network's input builder, then the preparation's inputs twice
        bt.prepare_only(prep)
        bt.propose(net)
        Trel, _, _, st = bt.retrack(None, n)
        assert np.all(st == 0)
        got.append(Trel.copy())
    assert np.allclose(got[0], expect, atol=2e-6), np.abs(got[0] - expect).max()
    for g in got[1:]:
        assert np.array_equal(g, got[0]), np.abs(g - got[0]).max()


@pytest.mark.parametrize("flag", range(8))
def test_augmentation_is_the_camera_symmetry(flag):
    """Mirroring the columns (rows) of both frames mirrors camera x (y):
    Trel -> S Trel S; swapping the frames inverts Trel. The augmented target
    must be the target of the transformed pose, and the augmented input the
    flipped / swapped tensor."""
    print(f"case cnn augment {flag}", flush=True)
    from recon_amd.hip import hip
    from recon_amd.synth import sample_motion
    cfg = default_config()
    net = PoseNet(4, config=cfg)
    n = 3
    rng = np.random.default_rng(flag)
    x = rng.normal(0, 1, (n, 60, 80, 4)).astype(np.float32)
    _, xb = bf16_bits(x)
    T = np.array([np.linalg.inv(sample_motion(100 + i, 10.0, 150.0)) for i in range(n)])
    y = net.pose_to_target(T)
    yi = net.pose_to_target(np.linalg.inv(T))
    bufs = [DeviceBuffer.from_array(a) for a in (xb, y, yi)]
    xo, yo = DeviceBuffer(xb.nbytes), DeviceBuffer(y.nbytes)
    net.augment(n, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, np.full(n, flag), xo.ptr, yo.ptr)
    hip().hipDeviceSynchronize()
    gx = xo.to_array(np.uint16, xb.size).reshape(xb.shape)
    gy = yo.to_array(np.float32, y.size).reshape(y.shape)
    ex = xb.copy()
    Te = T.copy()
    if flag & 4:
        ex = ex[..., [2, 3, 0, 1]]
        Te = np.linalg.inv(Te)
    if flag & 1:
        ex = ex[:, :, ::-1]
        S = np.diag([-1.0, 1, 1, 1]); Te = S @ Te @ S
    if flag & 2:
        ex = ex[:, ::-1]
        S = np.diag([1.0, -1, 1, 1]); Te = S @ Te @ S
    assert np.array_equal(gx, ex)
    assert np.allclose(gy, net.pose_to_target(Te), atol=1e-6)
