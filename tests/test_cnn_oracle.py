"""The CNN oracle (oracle/cnn.hpp) pinned on the CPU: its forward pass and
gradients against torch's own float64 convolution and autograd, its gradients
against finite differences of its own FP64 forward, and its bfloat16 rounding
against torch's round-to-nearest-even conversion."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import lib, ptr

SMALL = np.array([4, 8, 8, 12, 8, 8, 16], np.int32)


def n_params(ch, H, W):
    L = lib()
    L.or_cnn_num_params.restype = ctypes.c_size_t
    return int(L.or_cnn_num_params(ptr(ch), H, W))


def oracle_fb(ch, H, W, params, x, y, wn, round_=False, grad=True):
    L = lib()
    L.or_cnn_forward_backward.restype = ctypes.c_double
    B = x.shape[0]
    out = np.zeros(B * 6)
    g = np.zeros(n_params(ch, H, W)) if grad else None
    loss = L.or_cnn_forward_backward(ptr(ch), H, W, B, ptr(params), ptr(x), ptr(y), ptr(wn), int(round_), ptr(out),
                                     ptr(g) if grad else None)
    return loss, out.reshape(B, 6), g


def random_case(ch, H, W, B, seed):
    rng = np.random.default_rng(seed)
    P = n_params(ch, H, W)
    params = np.zeros(P, np.float32)
    o = 0
    for l in range(6):
        ci, co = ch[l], ch[l + 1]
        nw = co * 9 * ci
        params[o:o + nw] = rng.normal(0, np.sqrt(2.0 / (9 * ci)), nw)
        o += nw
        params[o:o + co] = rng.normal(0, 0.05, co)
        o += co
    params[o:o + 6 * ch[6]] = rng.normal(0, 0.2, 6 * ch[6])
    params[o + 6 * ch[6]:] = rng.normal(0, 0.1, 6)
    x = rng.normal(0, 1, (B, H, W, ch[0])).astype(np.float32)
    y = rng.normal(0, 1, (B, 6)).astype(np.float32)
    wn = (rng.random(B) > 0.2).astype(np.float32)
    wn[0] = 1.0
    return params, x, y, wn


def torch_reference(ch, H, W, params, x, y, wn):
    """The same network in torch float64 (NCHW), returning loss, outputs and
    the flat gradient in the oracle's layout."""
    p = torch.tensor(params.astype(np.float64), requires_grad=True)
    o = 0
    a = torch.tensor(x.astype(np.float64)).permute(0, 3, 1, 2)
    for l in range(6):
        ci, co = int(ch[l]), int(ch[l + 1])
        w = p[o:o + co * 9 * ci].view(co, 3, 3, ci).permute(0, 3, 1, 2)
        o += co * 9 * ci
        b = p[o:o + co]
        o += co
        a = torch.relu(torch.nn.functional.conv2d(a, w, b, stride=2, padding=1))
    h = a.mean(dim=(2, 3))
    Wfc = p[o:o + 6 * int(ch[6])].view(6, int(ch[6]))
    bfc = p[o + 6 * int(ch[6]):]
    out = h @ Wfc.T + bfc
    wt = torch.tensor(wn.astype(np.float64))
    loss = (wt[:, None] * (out - torch.tensor(y.astype(np.float64))) ** 2).sum() / max(1.0, float(wn.sum()))
    loss.backward()
    return float(loss.detach()), out.detach().numpy(), p.grad.numpy()


@pytest.mark.parametrize("H,W,seed", [(12, 16, 0), (15, 9, 1)])
def test_forward_and_gradients_match_torch_float64(H, W, seed):
    params, x, y, wn = random_case(SMALL, H, W, 3, seed)
    loss, out, g = oracle_fb(SMALL, H, W, params, x, y, wn)
    tl, tout, tg = torch_reference(SMALL, H, W, params, x, y, wn)
    assert np.allclose(out, tout, rtol=1e-10, atol=1e-12)
    assert abs(loss - tl) <= 1e-10 * abs(tl)
    assert np.allclose(g, tg, rtol=1e-9, atol=1e-12)


def test_gradients_match_finite_differences():
    H, W = 12, 16
    params, x, y, wn = random_case(SMALL, H, W, 2, 7)
    _, _, g = oracle_fb(SMALL, H, W, params, x, y, wn)
    rng = np.random.default_rng(3)
    p64 = params.astype(np.float64)
    idx = rng.choice(len(params), 60, replace=False)
    # params are FP32 in the oracle's interface: perturb by a step that FP32 holds
    for i in idx:
        h = max(1e-3 * abs(p64[i]), 1e-3)
        pp, pm = params.copy(), params.copy()
        pp[i] += np.float32(h); pm[i] -= np.float32(h)
        hp, hm = float(pp[i]) - p64[i], p64[i] - float(pm[i])
        lp = oracle_fb(SMALL, H, W, pp, x, y, wn, grad=False)[0]
        lm = oracle_fb(SMALL, H, W, pm, x, y, wn, grad=False)[0]
        fd = (lp - lm) / (hp + hm)
        assert abs(fd - g[i]) <= 2e-4 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_bf16_rounding_is_round_to_nearest_even():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 1, 5000), rng.normal(0, 1e-30, 100), rng.normal(0, 1e30, 100),
                        [0.0, -0.0, 1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, 65504.0, np.inf, -np.inf]]
                       ).astype(np.float32)
    out = np.zeros_like(x)
    lib().or_bf16_round(ptr(x), len(x), ptr(out))
    ref = torch.from_numpy(x).to(torch.bfloat16).float().numpy()
    assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))
