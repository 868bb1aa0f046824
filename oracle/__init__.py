"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP tracker. Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref.so")
_lib = None
_ref = None


def build(with_reference=False, force=False):
    flags = ["-B"] if force else []
    subprocess.run(["make", "-s"] + flags + ["-C", HERE, "all"], check=True)
    if with_reference and os.path.isdir("/root/reference/tracker"):
        subprocess.run(["make", "-s"] + flags + ["-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = ctypes.CDLL(LIB)
        _lib.or_tracker_create_f.restype = ctypes.c_void_p
        _lib.or_tracker_create_d.restype = ctypes.c_void_p
        for s in ("f", "d"):
            getattr(_lib, f"or_tracker_create_{s}").argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                                                 ctypes.c_void_p, ctypes.c_float]
            for name in ("destroy", "track", "fuse_current", "refresh_keyframe", "prepare"):
                getattr(_lib, f"or_tracker_{name}_{s}").argtypes = [ctypes.c_void_p]
            getattr(_lib, f"or_tracker_set_reference_{s}").argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
            getattr(_lib, f"or_tracker_set_current_{s}").argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
            getattr(_lib, f"or_tracker_set_reference_points_{s}").argtypes = [ctypes.c_void_p] * 5
            getattr(_lib, f"or_tracker_set_mode_{s}").argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_float]
            getattr(_lib, f"or_tracker_set_trel_{s}").argtypes = [ctypes.c_void_p, ctypes.c_void_p]
            getattr(_lib, f"or_tracker_get_{s}").argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 4
            getattr(_lib, f"or_tracker_selected_{s}").argtypes = [ctypes.c_void_p, ctypes.c_void_p]
            getattr(_lib, f"or_tracker_photo_system_{s}").argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
            getattr(_lib, f"or_tracker_icp_system_{s}").argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
            getattr(_lib, f"or_tracker_keyframe_{s}").argtypes = [ctypes.c_void_p] * 4
    return _lib


def ref_lib():
    """The reference's own expm.c / basic_math.cpp, built from its sources in
    oracle/_ref (None where the reference tree was not available to build it)."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_LIB):
            return None
        _ref = ctypes.CDLL(REF_LIB)
    return _ref


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


GEOMETRIC, PHOTOMETRIC, BIOBJECTIVE = 0, 1, 2


class Tracker:
    """The oracle tracker: dtype np.float32 (literal FP32 transcription) or
    np.float64 (FP64 oracle)."""

    def __init__(self, width, height, calib, mode=PHOTOMETRIC, levels=0, iters=(1, 3, 10, 5), nselect=8192,
                 mad_bins=64, robust=0, solver=0, lam=0.9, dtype=np.float64):
        self.L = lib()
        self.s = "f" if dtype == np.float32 else "d"
        self.dtype = dtype
        self.W, self.H = width, height
        self.calib = np.ascontiguousarray(calib, np.float32)
        prm = np.array([mode, levels] + list(iters) + [nselect, mad_bins, robust, solver], np.int32)
        self.h = getattr(self.L, f"or_tracker_create_{self.s}")(width, height, ptr(self.calib), ptr(prm), lam)

    def _f(self, name):
        return getattr(self.L, f"or_tracker_{name}_{self.s}")

    def __del__(self):
        try:
            self._f("destroy")(self.h)
        except Exception:
            pass

    def set_reference(self, gray, depth):
        g = np.ascontiguousarray(gray, np.uint8); d = np.ascontiguousarray(depth, np.uint16)
        self._f("set_reference")(self.h, ptr(g), ptr(d))

    def set_reference_points(self, gray, depth, xyz, normals):
        """Keyframe of a filtered reference: photometric from the frame, ICP
        cloud from ray-cast float4 maps (h, w, 4)."""
        g = np.ascontiguousarray(gray, np.uint8); d = np.ascontiguousarray(depth, np.uint16)
        x = np.ascontiguousarray(xyz, np.float32); n = np.ascontiguousarray(normals, np.float32)
        self._f("set_reference_points")(self.h, ptr(g), ptr(d), ptr(x), ptr(n))

    def set_current(self, gray, depth):
        g = np.ascontiguousarray(gray, np.uint8); d = np.ascontiguousarray(depth, np.uint16)
        self._f("set_current")(self.h, ptr(g), ptr(d))

    def track(self):
        self._f("track")(self.h)

    def set_mode(self, mode, lam):
        self._f("set_mode")(self.h, mode, lam)

    def set_trel(self, T):
        t = np.ascontiguousarray(T, np.float64).ravel()
        self._f("set_trel")(self.h, ptr(t))

    def poses(self):
        out = [np.zeros(16) for _ in range(4)]
        self._f("get")(self.h, *[ptr(o) for o in out])
        return [o.reshape(4, 4) for o in out]  # Trel, T, Tinc, global

    def selected(self, n):
        out = np.zeros(n, np.int32)
        self._f("selected")(self.h, ptr(out))
        return out

    def prepare(self):
        self._f("prepare")(self.h)

    def photo_system(self, layer):
        A = np.zeros(36); b = np.zeros(6)
        self._f("photo_system")(self.h, layer, ptr(A), ptr(b))
        return A.reshape(6, 6), b

    def icp_system(self, layer, it):
        A = np.zeros(36); b = np.zeros(6)
        n = self._f("icp_system")(self.h, layer, it, ptr(A), ptr(b))
        return A.reshape(6, 6), b, n

    def fuse_current(self):
        self._f("fuse_current")(self.h)

    def refresh_keyframe(self):
        self._f("refresh_keyframe")(self.h)

    def keyframe(self):
        z = np.zeros(self.W * self.H, self.dtype); g = np.zeros_like(z); w = np.zeros_like(z)
        self._f("keyframe")(self.h, ptr(z), ptr(g), ptr(w))
        return z, g, w


def call(name, *args):
    """Call a per-function oracle entry point, numpy arrays passed as pointers."""
    f = getattr(lib(), name)
    conv = []
    for a in args:
        if isinstance(a, np.ndarray):
            conv.append(ptr(a))
        elif isinstance(a, float):
            conv.append(ctypes.c_double(a))
        else:
            conv.append(a)
    return f(*conv)
