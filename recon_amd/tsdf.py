"""Truncated signed distance fusion and ray casting through the C ABI
(include/rt_tsdf.h): the voxel-grid keyframe model of the reference
(reconstruct/VoxelGrid.cpp, kernels/VoxelGrid_cuda.cu)."""
import ctypes

import numpy as np

from . import lib
from .hip import DeviceBuffer

_fp = ctypes.POINTER(ctypes.c_float)


class TsdfParams(ctypes.Structure):
    _fields_ = [("resolution", ctypes.c_int), ("origin", ctypes.c_float * 3), ("dim", ctypes.c_float * 3),
                ("width", ctypes.c_int), ("height", ctypes.c_int), ("min_dist", ctypes.c_float),
                ("max_dist", ctypes.c_float)]


def _check(rc, what):
    if rc != 0:
        msg = lib().rt_last_error()
        raise RuntimeError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def _mat(a, n):
    a = np.ascontiguousarray(np.asarray(a, np.float32).reshape(-1))
    if a.size != n:
        raise ValueError(f"expected {n} floats, got {a.size}")
    return a


def default_params(**kw):
    """The reference's grid (256^3, centre (0, 0.01, -1600.01), half-extent
    1500 mm, 320x240, 300..7000 mm), with fields overridden by keyword."""
    p = TsdfParams()
    lib().rt_tsdf_default_params(ctypes.byref(p))
    for k, v in kw.items():
        if k in ("origin", "dim"):
            getattr(p, k)[:] = [float(x) for x in v]
        else:
            setattr(p, k, v)
    return p


class VoxelGrid:
    """A TSDF grid on one device. Poses are camera -> world 4x4 (mm); K is
    the normalised 3x3 intrinsics (fx/w, cx/w, fy/h, cy/h). Updates are
    queued on the grid's stream; raycast() and read() return host arrays
    after waiting for it."""

    def __init__(self, params=None, device=0):
        self.params = params or default_params()
        h = ctypes.c_void_p()
        _check(lib().rt_tsdf_create(device, ctypes.byref(self.params), ctypes.byref(h)), "rt_tsdf_create")
        self.h = h
        w, hh = self.params.width, self.params.height
        self.res = self.params.resolution
        self._xyz = DeviceBuffer(w * hh * 16)
        self._nrm = DeviceBuffer(w * hh * 16)
        self._shade = DeviceBuffer(w * hh * 16)
        self._depth = None

    def close(self):
        if self.h:
            lib().rt_tsdf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self):
        return lib().rt_tsdf_stream(self.h)

    def synchronize(self):
        _check(lib().rt_tsdf_synchronize(self.h), "rt_tsdf_synchronize")

    def reset(self):
        _check(lib().rt_tsdf_reset(self.h), "rt_tsdf_reset")

    def set_depth(self, depth):
        """Distance map for the next update: a host (h, w) uint16 array in mm,
        or a host (h, w, 4) float32 RGBD image in the reference's layout (w =
        normalised depth, rows bottom-up)."""
        a = np.ascontiguousarray(depth)
        w, h = self.params.width, self.params.height
        self.synchronize()  # the staging buffer may still be read by queued work
        if self._depth is None or self._depth.nbytes < a.nbytes:
            self._depth = DeviceBuffer(max(a.nbytes, w * h * 16))
        if a.dtype == np.uint16 and a.shape == (h, w):
            self._depth.upload(a)
            _check(lib().rt_tsdf_set_depth_u16(self.h, self._depth.ptr), "rt_tsdf_set_depth_u16")
        elif a.dtype == np.float32 and a.shape == (h, w, 4):
            self._depth.upload(a)
            _check(lib().rt_tsdf_set_depth_rgbd(self.h, self._depth.ptr), "rt_tsdf_set_depth_rgbd")
        else:
            raise ValueError(f"depth must be uint16 ({h}, {w}) or float32 ({h}, {w}, 4), got {a.dtype} {a.shape}")

    def update(self, T, K):
        _check(lib().rt_tsdf_update(self.h, _mat(T, 16).ctypes.data_as(_fp), _mat(K, 9).ctypes.data_as(_fp)),
               "rt_tsdf_update")

    def integrate(self, depth, T, K):
        self.set_depth(depth)
        self.update(T, K)

    def resample_from(self, src, T):
        """This grid := src resampled under pose T (camera -> world of the new
        anchor); both grids have the same geometry."""
        _check(lib().rt_tsdf_resample(self.h, src.h, _mat(T, 16).ctypes.data_as(_fp)), "rt_tsdf_resample")

    def raycast(self, T, K, refine=True):
        """(xyz, normals, shade) host arrays, each (h, w, 4) float32."""
        _check(lib().rt_tsdf_raycast(self.h, _mat(T, 16).ctypes.data_as(_fp), _mat(K, 9).ctypes.data_as(_fp),
                                     int(refine), self._xyz.ptr, self._nrm.ptr, self._shade.ptr), "rt_tsdf_raycast")
        self.synchronize()
        w, h = self.params.width, self.params.height
        n = w * h * 4
        return tuple(b.to_array(np.float32, n).reshape(h, w, 4) for b in (self._xyz, self._nrm, self._shade))

    def read(self):
        """The grid as (res, res, res, 2) float32 [z, y, x, (distance, weight)]."""
        r = self.res
        out = np.empty((r, r, r, 2), np.float32)
        _check(lib().rt_tsdf_read(self.h, out.ctypes.data, out.nbytes), "rt_tsdf_read")
        return out
