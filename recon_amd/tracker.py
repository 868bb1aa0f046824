"""Python mirror of the reference's PhotoTracker API over the HIP C ABI.

Method names follow tracker/tracker/phototracker.h:33-82 (snake_case); every
call goes to librecon_amd.so and raises RuntimeError on failure.
"""
import ctypes

import numpy as np

from . import (BIOBJECTIVE, GEOMETRIC, PHOTOMETRIC, RtPair, RtPairResult, RtParams, check, lib)

DBG = dict(REF_GRAY=0, CUR_GRAY=1, REF_Z=2, CUR_Z=3, GRAD_MAG=4, GRAD_HIST=5, SELECTED=6, POINTS=7, JACOBIAN=8,
           RESIDUAL=9, TREL=10, REF_CLOUD=11, CUR_CLOUD=12, ICP_SELECTED=13, ICP_COUNT=14, ICP_R2=15, ICP_VAR=16,
           SYSTEM=17, KF_WEIGHT=18, MAD_HIST=19, DEPTH_BLOCKS=20)


def default_params(**kw):
    p = RtParams()
    lib().rt_default_params(ctypes.byref(p))
    for k, v in kw.items():
        if k == "iters":
            for i, x in enumerate(v):
                p.iters[i] = x
        elif k == "lam":
            p.lambda_ = v
        else:
            setattr(p, k, v)
    return p


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


class PhotoTracker:
    """One tracker on one device (initializePhotoTracker + allocMemory)."""

    def __init__(self, width, height, calib, device=0, **params):
        self.L = lib()
        self.W, self.H = width, height
        self.params = default_params(**params)
        self.calib = _f32(calib)
        h = ctypes.c_void_p()
        check(self.L.rt_create(device, width, height, self.calib.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                               ctypes.byref(self.params), ctypes.byref(h)), "rt_create")
        self.h = h
        self.n_select = self.params.n_select

    def close(self):
        if getattr(self, "h", None):
            self.L.rt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- frames
    def set_current_image(self, gray, depth, on_device=False):
        if on_device:
            return check(self.L.rt_set_current(self.h, gray, depth, 1), "rt_set_current")
        g = np.ascontiguousarray(gray, np.uint8); d = np.ascontiguousarray(depth, np.uint16)
        check(self.L.rt_set_current(self.h, g.ctypes.data, d.ctypes.data, 0), "rt_set_current")
        self.L.rt_synchronize(self.h)

    def set_current_rgbd(self, rgbd):
        a = _f32(rgbd)
        check(self.L.rt_set_current_rgbd(self.h, a.ctypes.data, 0), "rt_set_current_rgbd")
        self.L.rt_synchronize(self.h)

    def set_reference_points(self, gray, depth, mode=PHOTOMETRIC, lam=0.0, force=False, on_device=False):
        if on_device:
            return check(self.L.rt_set_reference(self.h, gray, depth, 1, mode, lam, int(force)), "rt_set_reference")
        g = np.ascontiguousarray(gray, np.uint8); d = np.ascontiguousarray(depth, np.uint16)
        r = check(self.L.rt_set_reference(self.h, g.ctypes.data, d.ctypes.data, 0, mode, lam, int(force)),
                  "rt_set_reference")
        self.L.rt_synchronize(self.h)
        return bool(r)

    def set_reference_from_points(self, xyz, normals, gray, depth, mode=PHOTOMETRIC, lam=0.0, force=False,
                                  on_device=False):
        """setReferencePoints with ray-cast maps: xyz and normals are device
        pointers to float4 images of the tracker's size (e.g. a VoxelGrid's
        ray cast); the frame (gray, depth) is host or device data."""
        if on_device:
            return bool(check(self.L.rt_set_reference_points(self.h, xyz, normals, gray, depth, 1, mode, lam, int(force)),
                              "rt_set_reference_points"))
        g = np.ascontiguousarray(gray, np.uint8); d = np.ascontiguousarray(depth, np.uint16)
        r = check(self.L.rt_set_reference_points(self.h, xyz, normals, g.ctypes.data, d.ctypes.data, 0, mode, lam,
                                                 int(force)), "rt_set_reference_points")
        self.L.rt_synchronize(self.h)
        return bool(r)

    def update_error_image(self):
        check(self.L.rt_update_error_image(self.h), "rt_update_error_image")

    def _image(self, fn, level):
        w, h = (self.W, self.H) if self.get_mode() == PHOTOMETRIC else (self.W >> level, self.H >> level)
        out = np.empty((h, w), np.float32)
        check(fn(self.h, level, out.ctypes.data, out.size), fn.__name__)
        return out

    def get_debug_image(self, level=0):
        """getDebugImage1C: the error image (host array)."""
        return self._image(self.L.rt_get_debug_image, level)

    def get_ref_image(self, level=0):
        """getRefImage1C: the keyframe image for display (host array)."""
        return self._image(self.L.rt_get_ref_image, level)

    def set_filtered_reference(self, flag):
        check(self.L.rt_set_filtered_reference(self.h, int(flag)), "rt_set_filtered_reference")

    def get_filtered_reference(self):
        return bool(check(self.L.rt_get_filtered_reference(self.h), "rt_get_filtered_reference"))

    def set_reference_rgbd(self, rgbd, mode=PHOTOMETRIC, lam=0.0, force=False):
        a = _f32(rgbd)
        r = check(self.L.rt_set_reference_rgbd(self.h, a.ctypes.data, 0, mode, lam, int(force)), "rt_set_reference_rgbd")
        self.L.rt_synchronize(self.h)
        return bool(r)

    def track(self):
        return bool(check(self.L.rt_track(self.h), "rt_track"))

    def track_from(self, T_init):
        """Track the current frame starting from T_init (current camera relative
        to the keyframe)."""
        t = np.ascontiguousarray(T_init, np.float64).ravel()
        return bool(check(self.L.rt_track_from(self.h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                          "rt_track_from"))

    def residual_rms(self):
        r = ctypes.c_float()
        check(self.L.rt_get_residual_rms(self.h, ctypes.byref(r)), "rt_get_residual_rms")
        return float(r.value)

    # ---- poses
    def get_global_pose(self):
        T = np.zeros(16, np.float64)
        check(self.L.rt_get_global_pose(self.h, T.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return T.reshape(4, 4)

    def get_relative_pose(self):
        T = np.zeros(16, np.float32)
        check(self.L.rt_get_relative_pose(self.h, T.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return T.reshape(4, 4)

    def get_increment(self):
        T = np.zeros(16, np.float64)
        check(self.L.rt_get_increment(self.h, T.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return T.reshape(4, 4)

    def trel(self):
        return self.debug_read("TREL", 0, np.float32, 16).reshape(4, 4)

    # ---- settings
    def reset_tracking(self, mode=GEOMETRIC):
        check(self.L.rt_reset_tracking(self.h, mode))

    def clear_pose(self):
        check(self.L.rt_clear_pose(self.h))

    def reset_to_transform(self, T):
        t = np.ascontiguousarray(T, np.float64).ravel()
        check(self.L.rt_reset_to_transform(self.h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    def set_mode(self, mode):
        check(self.L.rt_set_mode(self.h, mode))

    def get_mode(self):
        return check(self.L.rt_get_mode(self.h))

    def set_lambda(self, lam):
        check(self.L.rt_set_lambda(self.h, lam))

    def get_lambda(self):
        return self.L.rt_get_lambda(self.h)

    def set_iteration_counts(self, n):
        a = (ctypes.c_int * 4)(*list(n) + [0] * (4 - len(n)))
        check(self.L.rt_set_iteration_counts(self.h, a))

    def get_iteration_counts(self):
        a = (ctypes.c_int * 4)()
        check(self.L.rt_get_iteration_counts(self.h, a))
        return list(a)

    def set_reference_updating(self, flag):
        check(self.L.rt_set_reference_updating(self.h, int(flag)))

    def set_incremental_mode(self, flag):
        check(self.L.rt_set_incremental_mode(self.h, int(flag)))

    def signal_calibration_update(self, calib):
        c = _f32(calib)
        check(self.L.rt_signal_calibration_update(self.h, c.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))

    def get_layers(self):
        return check(self.L.rt_get_layers(self.h))

    def num_reference_updates(self):
        return check(self.L.rt_num_reference_updates(self.h))

    def list_selected(self):
        n = self.n_select
        cur = np.zeros(2 * n, np.float32); ref = np.zeros(2 * n, np.float32); w = np.zeros(n, np.float32)
        cnt = ctypes.c_int()
        fp = ctypes.POINTER(ctypes.c_float)
        check(self.L.rt_list_selected(self.h, cur.ctypes.data_as(fp), ref.ctypes.data_as(fp), w.ctypes.data_as(fp), n,
                                      ctypes.byref(cnt)))
        k = cnt.value
        return cur[:2 * k].reshape(-1, 2), ref[:2 * k].reshape(-1, 2), w[:k]

    def fuse_current(self):
        check(self.L.rt_fuse_current(self.h))

    def refresh_keyframe(self):
        check(self.L.rt_refresh_keyframe(self.h))

    def synchronize(self):
        check(self.L.rt_synchronize(self.h))

    def track_sequence(self, gray_dev_ptr, depth_dev_ptr, n, keyframe_interval=10, fuse_frames=5):
        poses = np.zeros((n, 16), np.float64)
        check(self.L.rt_track_sequence(self.h, gray_dev_ptr, depth_dev_ptr, n, keyframe_interval, fuse_frames,
                                       poses.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "rt_track_sequence")
        return poses.reshape(n, 4, 4)

    # ---- inspection
    def debug_read(self, what, level, dtype, count):
        out = np.zeros(count, dtype)
        check(self.L.rt_debug_read(self.h, DBG[what], level, out.ctypes.data, out.nbytes), f"debug_read {what}")
        return out

    def debug_set_trel(self, T):
        t = _f32(T).ravel()
        check(self.L.rt_debug_set_trel(self.h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))

    def debug_prepare(self):
        check(self.L.rt_debug_prepare(self.h), "rt_debug_prepare")

    def debug_linearize(self, level, it=0):
        check(self.L.rt_debug_linearize(self.h, level, it), "rt_debug_linearize")
        s = self.debug_read("SYSTEM", 0, np.float64, 54)
        return _unpack27(s[:27]), _unpack27(s[27:])


def _unpack27(s):
    A = np.zeros((6, 6))
    k = 0
    for r in range(6):
        for c in range(r, 6):
            A[r, c] = A[c, r] = s[k]
            k += 1
    return A, s[21:27].copy()


class BatchTracker:
    """Many frame pairs tracked by one persistent launch (rt_batch_*)."""

    def __init__(self, max_pairs, max_width, max_height, device=0, **params):
        self.L = lib()
        self.params = default_params(**params)
        h = ctypes.c_void_p()
        check(self.L.rt_batch_create(device, max_pairs, max_width, max_height, ctypes.byref(self.params),
                                     ctypes.byref(h)), "rt_batch_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.rt_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def track(self, pairs, init=None):
        """pairs: list of dicts with device pointers ref_gray, ref_depth, cur_gray,
        cur_depth (ints), width, height, calib (np.float32[96]); or the result
        of prepare()."""
        if isinstance(pairs, tuple) and init is not None:
            # a prepared array carries its own initial poses: prepare, then track from init
            self.prepare_only(pairs)
            return self.retrack(np.stack([np.asarray(t, np.float64).reshape(4, 4) for t in init]), pairs[1])
        prep = pairs if isinstance(pairs, tuple) else self.prepare(pairs, init)
        arr, n, out, _keep = prep
        check(self.L.rt_batch_track(self.h, n, arr, out), "rt_batch_track")
        return _results(out)

    def prepare(self, pairs, init=None):
        """The ctypes pair array for track(), built once for repeated calls."""
        n = len(pairs)
        arr = (RtPair * n)()
        keep = []
        for i, p in enumerate(pairs):
            c = _f32(p["calib"])
            keep.append(c)
            arr[i].ref_gray = p["ref_gray"]; arr[i].ref_depth = p["ref_depth"]
            arr[i].cur_gray = p["cur_gray"]; arr[i].cur_depth = p["cur_depth"]
            arr[i].width = p["width"]; arr[i].height = p["height"]
            arr[i].calib = c.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
            if init is not None and init[i] is not None:
                t = np.ascontiguousarray(init[i], np.float64).ravel()
                keep.append(t)
                arr[i].init_T = t.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        out = (RtPairResult * n)()
        return (arr, n, out, keep)

    def prepare_only(self, prepared):
        """Keyframe preparation of the pairs of prepare() without tracking."""
        arr, n, _out, _keep = prepared
        check(self.L.rt_batch_prepare(self.h, n, arr), "rt_batch_prepare")

    def debug_read(self, pair, what, level, dtype, count):
        out = np.zeros(count, dtype)
        check(self.L.rt_batch_debug_read(self.h, pair, DBG[what], level, out.ctypes.data, out.nbytes),
              "rt_batch_debug_read")
        return out

    def propose(self, net):
        """Initial poses of the prepared pairs from the pose network (PoseNet)."""
        check(self.L.rt_batch_propose(self.h, net.h), "rt_batch_propose")

    def retrack(self, init=None, n=None):
        out = (RtPairResult * n)()
        ip = None
        if init is not None:
            t = np.ascontiguousarray(init, np.float64).ravel()
            ip = t.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        check(self.L.rt_batch_retrack(self.h, ip, out), "rt_batch_retrack")
        return _results(out)

    def results(self, n):
        """Results of the last tracking launch enqueued without output
        (retrack_async), after waiting for it."""
        out = (RtPairResult * n)()
        check(self.L.rt_batch_results(self.h, out), "rt_batch_results")
        return _results(out)

    def retrack_async(self):
        """Enqueue a tracking launch from the initial poses on the device
        (the proposals) without waiting; results() collects it."""
        check(self.L.rt_batch_retrack(self.h, None, None), "rt_batch_retrack")


_RESULT_DTYPE = np.dtype([("Trel", np.float64, (4, 4)), ("T", np.float64, (4, 4)), ("rms", np.float32),
                          ("status", np.int32)])


def _results(out):
    a = np.frombuffer(out, dtype=_RESULT_DTYPE)
    return a["Trel"].copy(), a["T"].copy(), a["rms"].copy(), a["status"].copy()
