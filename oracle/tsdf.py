"""numpy face of the TSDF oracle (tsdf.hpp in liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker of the HIP TSDF kernels.
"""
import ctypes

import numpy as np

from . import lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def _prm(res, w, h, origin, dim, min_dist, max_dist):
    prm = np.array([res, w, h], np.int32)
    box = np.array(list(origin) + list(dim) + [min_dist, max_dist], np.float32)
    return prm, box


class Grid:
    """A (res, res, res, 2) float32 grid of (distance, weight), x fastest."""

    def __init__(self, res, w, h, origin, dim, min_dist=300.0, max_dist=7000.0):
        self.res, self.w, self.h = res, w, h
        self.prm, self.box = _prm(res, w, h, origin, dim, min_dist, max_dist)
        self.data = np.empty((res, res, res, 2), np.float32)
        self.reset()

    def reset(self):
        self.data[..., 0] = 1.0
        self.data[..., 1] = 0.0

    def dist_from_rgbd(self, rgbd):
        rgbd = np.ascontiguousarray(rgbd, np.float32)
        assert rgbd.shape == (self.h, self.w, 4)
        out = np.empty((self.h, self.w), np.float32)
        lib().or_tsdf_dist_rgbd(_p(self.prm), _p(self.box), _p(rgbd), _p(out))
        return out

    def update(self, dist, T, K):
        dist = np.ascontiguousarray(dist, np.float32)
        assert dist.shape == (self.h, self.w)
        T = np.ascontiguousarray(np.asarray(T, np.float32).reshape(16))
        K = np.ascontiguousarray(np.asarray(K, np.float32).reshape(9))
        lib().or_tsdf_update(_p(self.prm), _p(self.box), _p(self.data), _p(dist), _p(T), _p(K))

    def resample_from(self, src, T):
        """This grid := src resampled under pose T (the grids share their geometry)."""
        T = np.ascontiguousarray(np.asarray(T, np.float32).reshape(16))
        lib().or_tsdf_resample(_p(self.prm), _p(self.box), _p(self.data), _p(src.data), _p(T))

    def raycast(self, T, K, refine=True):
        T = np.ascontiguousarray(np.asarray(T, np.float32).reshape(16))
        K = np.ascontiguousarray(np.asarray(K, np.float32).reshape(9))
        xyz = np.empty((self.h, self.w, 4), np.float32)
        nrm = np.empty_like(xyz)
        shade = np.empty_like(xyz)
        lib().or_tsdf_raycast(_p(self.prm), _p(self.box), _p(self.data), _p(T), _p(K), ctypes.c_int(int(refine)),
                              _p(xyz), _p(nrm), _p(shade))
        return xyz, nrm, shade
