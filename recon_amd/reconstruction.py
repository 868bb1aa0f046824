"""Keyframe tracking against the fused TSDF model: the per-frame sequence of
the reference application (updateVoxelGrid, rgbdscan/src/main.cpp:923-986,
with the relative TSDF of getTSDFPose, :407-426, and resampleVoxelGrid,
:884-895), on one GPU: the previous frame is fused into the grid at its pose
relative to the keyframe, the grid is ray-cast from that pose, and the ray
cast becomes the ICP keyframe whenever the tracker's keyframe-switch rule asks
for a new one (setReferencePoints with the filtered reference); on a switch
the grid is resampled into the new keyframe's frame and the tracker restarts
from the last frame-to-frame increment."""
import numpy as np

from . import GEOMETRIC
from .tracker import PhotoTracker
from .tsdf import VoxelGrid, default_params


def normalized_k(calib, width, height):
    """getNormalizedK (rgbdscan/src/main.cpp:878-882): the depth camera's
    intrinsics KL of the calibration block divided by the image size."""
    K = np.array(calib[43:52], np.float32).copy()
    K[0] /= np.float32(width); K[1] = 0; K[2] /= np.float32(width)
    K[3] = 0; K[4] /= np.float32(height); K[5] /= np.float32(height)
    return K


class FusedKeyframeTracker:
    """frame(gray, depth) per frame (host arrays, 8-bit gray and 16-bit mm);
    returns the camera pose in the first frame's coordinates (camera ->
    world, float64 4x4). grid: rt_tsdf parameters (default: the reference's
    grid at this image size); relative: re-anchor the grid at each keyframe
    (the reference's relativeTSDF) instead of keeping it in the first frame's
    coordinates."""

    def __init__(self, width, height, calib, mode=GEOMETRIC, lam=0.5, grid=None, relative=True, device=0,
                 **tracker_params):
        self.W, self.H = width, height
        self.mode, self.lam, self.relative = mode, lam, relative
        self.tr = PhotoTracker(width, height, calib, device=device, mode=mode, **tracker_params)
        self.tr.set_filtered_reference(True)
        self.tr.set_reference_updating(True)
        p = grid if grid is not None else default_params(width=width, height=height)
        if p.width != width or p.height != height:
            raise ValueError("the grid's image size must be the tracker's")
        self.grids = [VoxelGrid(p, device), VoxelGrid(p, device) if relative else None]
        self.active = 0
        self.K = normalized_k(calib, width, height)
        self.prev = None
        self.prev_fused = False
        self.T_grid = np.eye(4, dtype=np.float32)  # pose of the previous frame in the grid's coordinates
        self.switches = 0

    @property
    def grid(self):
        return self.grids[self.active]

    def _reference(self, gray, depth, T, force):
        g = self.grid
        if not self.prev_fused:
            g.integrate(np.ascontiguousarray(depth, np.uint16), T, self.K)
        g.raycast(T, self.K, True)
        return self.tr.set_reference_from_points(g._xyz.ptr, g._nrm.ptr, gray, depth, mode=self.mode, lam=self.lam,
                                                 force=force)

    def frame(self, gray, depth):
        if self.prev is None:
            self._reference(gray, depth, np.eye(4, dtype=np.float32), True)
            self.prev, self.prev_fused = (gray, depth), True
            return np.eye(4)
        g0, d0 = self.prev
        # the previous frame at its pose relative to the keyframe (getTSDFPose)
        T = self.tr.get_relative_pose().astype(np.float32) if self.relative else self.T_grid
        if self._reference(g0, d0, T, False):
            self.switches += 1
            if self.relative:
                # resampleVoxelGrid: the grid moves into the new keyframe's frame
                other = 1 - self.active
                if self.grids[other] is None:
                    self.grids[other] = VoxelGrid(self.grid.params)
                self.grids[other].resample_from(self.grid, T)
                self.active = other
                self.tr.reset_to_transform(self.tr.get_increment())
        self.tr.set_current_image(gray, depth)
        self.tr.track()
        self.prev, self.prev_fused = (gray, depth), False
        if not self.relative:
            self.T_grid = self.tr.get_global_pose().astype(np.float32)
        return self.tr.get_global_pose()
