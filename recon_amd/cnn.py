"""Pose-proposal network (include/rt_cnn.h) from Python: parameters, inputs,
forward pass, gradients and training steps, all on the GPU through the C ABI."""
import ctypes
import os

import numpy as np

from . import RtCnnConfig, RtPair, check, lib

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "data", "pose_cnn.npz")


def default_config(**kw):
    c = RtCnnConfig()
    lib().rt_cnn_default_config(ctypes.byref(c))
    for k, v in kw.items():
        if k == "channels":
            for i, x in enumerate(v):
                c.channels[i] = int(x)
        else:
            setattr(c, k, v)
    return c


def layout(cfg):
    """Offsets of the flat parameter vector: [(w_off, b_off, cout, cin)] * 6, fc_w, fc_b, total."""
    ch = list(cfg.channels)
    o, layers = 0, []
    for l in range(6):
        w = o
        o += ch[l + 1] * 9 * ch[l]
        layers.append((w, o, ch[l + 1], ch[l]))
        o += ch[l + 1]
    fc_w = o
    o += 6 * ch[6]
    return layers, fc_w, o, o + 6


def init_params(cfg, seed=0):
    """He-normal conv weights, zero biases, small dense layer."""
    rng = np.random.default_rng(seed)
    layers, fc_w, fc_b, total = layout(cfg)
    p = np.zeros(total, np.float32)
    for (w, b, co, ci) in layers:
        p[w:b] = rng.normal(0.0, np.sqrt(2.0 / (9 * ci)), b - w)
    p[fc_w:fc_b] = rng.normal(0.0, 0.01, fc_b - fc_w)
    return p


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class PoseNet:
    def __init__(self, max_batch, device=0, config=None):
        self.L = lib()
        self.cfg = config if config is not None else default_config()
        h = ctypes.c_void_p()
        check(self.L.rt_cnn_create(device, ctypes.byref(self.cfg), max_batch, ctypes.byref(h)), "rt_cnn_create")
        self.h = h
        self.max_batch = max_batch
        self.n_params = int(self.L.rt_cnn_num_params(self.h))

    def close(self):
        if getattr(self, "h", None):
            self.L.rt_cnn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def input_elems(self):
        return self.cfg.in_width * self.cfg.in_height * self.cfg.channels[0]

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        check(self.L.rt_cnn_set_params(self.h, _fp(p)), "rt_cnn_set_params")

    def get_params(self):
        p = np.zeros(self.n_params, np.float32)
        check(self.L.rt_cnn_get_params(self.h, _fp(p)), "rt_cnn_get_params")
        return p

    def save(self, path):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.savez(path, params=self.get_params(), channels=np.array(list(self.cfg.channels), np.int32),
                 size=np.array([self.cfg.in_width, self.cfg.in_height], np.int32),
                 scales=np.array([self.cfg.rot_scale_deg, self.cfg.trans_scale_mm], np.float32))

    @classmethod
    def load(cls, path=DEFAULT_WEIGHTS, max_batch=4096, device=0):
        with np.load(path, allow_pickle=False) as z:
            cfg = default_config(channels=z["channels"].tolist(), in_width=int(z["size"][0]), in_height=int(z["size"][1]),
                                 rot_scale_deg=float(z["scales"][0]), trans_scale_mm=float(z["scales"][1]))
            net = cls(max_batch, device, cfg)
            net.set_params(z["params"])
        return net

    def make_inputs(self, pair_array, n, x_ptr):
        check(self.L.rt_cnn_make_inputs(self.h, n, pair_array, x_ptr), "rt_cnn_make_inputs")

    def forward(self, n, x_ptr, out_ptr):
        check(self.L.rt_cnn_forward(self.h, n, x_ptr, out_ptr), "rt_cnn_forward")

    def loss_grad(self, n, x_ptr, y_ptr, w_ptr):
        loss = ctypes.c_float()
        out = np.zeros(6 * n, np.float32)
        g = np.zeros(self.n_params, np.float32)
        check(self.L.rt_cnn_loss_grad(self.h, n, x_ptr, y_ptr, w_ptr, ctypes.byref(loss), _fp(out), _fp(g)),
              "rt_cnn_loss_grad")
        return float(loss.value), out.reshape(n, 6), g

    def train_step(self, n, x_ptr, y_ptr, w_ptr, lr, want_loss=True):
        loss = ctypes.c_float()
        check(self.L.rt_cnn_train_step(self.h, n, x_ptr, y_ptr, w_ptr, ctypes.c_float(lr),
                                       ctypes.byref(loss) if want_loss else None), "rt_cnn_train_step")
        return float(loss.value)

    def augment(self, n, x_ptr, y_ptr, y_inv_ptr, flags, x_out_ptr, y_out_ptr):
        f = np.ascontiguousarray(flags, np.int32)
        check(self.L.rt_cnn_augment(self.h, n, x_ptr, y_ptr, y_inv_ptr, f.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                    x_out_ptr, y_out_ptr), "rt_cnn_augment")

    def pose_to_target(self, Trel):
        return pose_to_target(self.cfg, Trel)

    def target_to_pose(self, y):
        return target_to_pose(self.cfg, y)


def pose_to_target(cfg, Trel):
    """Network targets (scaled rotation vector, scaled translation) of Trel poses."""
    T = np.ascontiguousarray(Trel, np.float64).reshape(-1, 16)
    y = np.zeros((len(T), 6), np.float32)
    check(lib().rt_cnn_pose_to_target(ctypes.byref(cfg), len(T), T.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), _fp(y)),
          "rt_cnn_pose_to_target")
    return y


def target_to_pose(cfg, y):
    y = np.ascontiguousarray(y, np.float32).reshape(-1, 6)
    T = np.zeros((len(y), 4, 4))
    check(lib().rt_cnn_target_to_pose(ctypes.byref(cfg), len(y), _fp(y), T.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
          "rt_cnn_target_to_pose")
    return T


def pair_array(pairs):
    """ctypes RtPair array from dicts of device pointers (no calib needed)."""
    arr = (RtPair * len(pairs))()
    for i, p in enumerate(pairs):
        arr[i].ref_gray = p["ref_gray"]; arr[i].ref_depth = p["ref_depth"]
        arr[i].cur_gray = p["cur_gray"]; arr[i].cur_depth = p["cur_depth"]
        arr[i].width = p["width"]; arr[i].height = p["height"]
    return arr
