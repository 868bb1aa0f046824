"""MI355X-native RGB-D keyframe tracker (photometric, geometric, bi-objective).

The compute path is the HIP library ``librecon_amd.so`` built in-tree for
gfx950 (``python -m recon_amd.build`` or ``__graft_entry__.build()``); this
package only binds its C ABI (``include/rt_tracker.h``) with ctypes. There is no
CPU fallback: loading fails loudly when the library is missing.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librecon_amd.so")

_lib = None
_synth = None

c_int, c_float, c_double, c_void_p, c_size_t = ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_void_p, ctypes.c_size_t
P = ctypes.POINTER


class RtParams(ctypes.Structure):
    _fields_ = [
        ("mode", c_int), ("lambda_", c_float), ("levels", c_int), ("iters", c_int * 4), ("n_select", c_int),
        ("mad_bins", c_int), ("robust", c_int), ("solver", c_int), ("pose_dist_threshold", c_float),
        ("pose_angle_threshold", c_float), ("incremental", c_int), ("reference_updating", c_int),
    ]


class RtPair(ctypes.Structure):
    _fields_ = [
        ("ref_gray", c_void_p), ("ref_depth", c_void_p), ("cur_gray", c_void_p), ("cur_depth", c_void_p),
        ("width", c_int), ("height", c_int), ("calib", P(c_float)), ("init_T", P(c_double)),
    ]


class RtPairResult(ctypes.Structure):
    _fields_ = [("Trel", c_double * 16), ("T", c_double * 16), ("rms", c_float), ("status", c_int)]


class RtCnnConfig(ctypes.Structure):
    _fields_ = [("in_width", c_int), ("in_height", c_int), ("channels", c_int * 7), ("rot_scale_deg", c_float),
                ("trans_scale_mm", c_float)]


# name -> (restype, argtypes); the exported symbols of include/rt_*.h
SIGNATURES = {
    "rt_default_params": (None, [P(RtParams)]),
    "rt_calib_pinhole": (c_int, [c_float, c_float, c_float, c_float, c_float, c_float, P(c_float)]),
    "rt_create": (c_int, [c_int, c_int, c_int, P(c_float), P(RtParams), P(c_void_p)]),
    "rt_destroy": (None, [c_void_p]),
    "rt_set_current": (c_int, [c_void_p, c_void_p, c_void_p, c_int]),
    "rt_set_current_rgbd": (c_int, [c_void_p, c_void_p, c_int]),
    "rt_set_reference": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_float, c_int]),
    "rt_set_reference_rgbd": (c_int, [c_void_p, c_void_p, c_int, c_int, c_float, c_int]),
    "rt_set_filtered_reference": (c_int, [c_void_p, c_int]),
    "rt_update_error_image": (c_int, [c_void_p]),
    "rt_get_debug_image": (c_int, [c_void_p, c_int, c_void_p, c_size_t]),
    "rt_get_ref_image": (c_int, [c_void_p, c_int, c_void_p, c_size_t]),
    "rt_get_filtered_reference": (c_int, [c_void_p]),
    "rt_set_reference_points": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_float, c_int]),
    "rt_set_reference_points_rgbd": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_float, c_int]),
    "rt_track": (c_int, [c_void_p]),
    "rt_get_global_pose": (c_int, [c_void_p, P(c_double)]),
    "rt_track_from": (c_int, [c_void_p, P(c_double)]),
    "rt_get_residual_rms": (c_int, [c_void_p, P(c_float)]),
    "rt_get_relative_pose": (c_int, [c_void_p, P(c_float)]),
    "rt_get_increment": (c_int, [c_void_p, P(c_double)]),
    "rt_reset_tracking": (c_int, [c_void_p, c_int]),
    "rt_clear_pose": (c_int, [c_void_p]),
    "rt_reset_to_transform": (c_int, [c_void_p, P(c_double)]),
    "rt_set_mode": (c_int, [c_void_p, c_int]),
    "rt_get_mode": (c_int, [c_void_p]),
    "rt_set_lambda": (c_int, [c_void_p, c_float]),
    "rt_get_lambda": (c_float, [c_void_p]),
    "rt_set_iteration_counts": (c_int, [c_void_p, P(c_int)]),
    "rt_get_iteration_counts": (c_int, [c_void_p, P(c_int)]),
    "rt_set_reference_updating": (c_int, [c_void_p, c_int]),
    "rt_set_incremental_mode": (c_int, [c_void_p, c_int]),
    "rt_signal_calibration_update": (c_int, [c_void_p, P(c_float)]),
    "rt_get_layers": (c_int, [c_void_p]),
    "rt_num_reference_updates": (c_int, [c_void_p]),
    "rt_list_selected": (c_int, [c_void_p, P(c_float), P(c_float), P(c_float), c_int, P(c_int)]),
    "rt_fuse_current": (c_int, [c_void_p]),
    "rt_refresh_keyframe": (c_int, [c_void_p]),
    "rt_stream": (c_void_p, [c_void_p]),
    "rt_synchronize": (c_int, [c_void_p]),
    "rt_track_sequence": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, P(c_double)]),
    "rt_batch_create": (c_int, [c_int, c_int, c_int, c_int, P(RtParams), P(c_void_p)]),
    "rt_batch_destroy": (None, [c_void_p]),
    "rt_batch_track": (c_int, [c_void_p, c_int, P(RtPair), P(RtPairResult)]),
    "rt_batch_retrack": (c_int, [c_void_p, P(c_double), P(RtPairResult)]),
    "rt_batch_results": (c_int, [c_void_p, P(RtPairResult)]),
    "rt_batch_stream": (c_void_p, [c_void_p]),
    "rt_batch_variant_info": (c_int, [c_int, P(c_int)]),
    "rt_batch_track_noalloc": (c_int, [c_int, c_int, P(RtPair), P(RtParams), c_void_p, P(RtPairResult)]),
    "rt_batch_layout": (c_int, [P(RtParams), c_int, c_int, c_int, P(RtPair), P(ctypes.c_int64), P(c_int), P(c_int)]),
    "rt_batch_debug_read": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_size_t]),
    "rt_debug_read": (c_int, [c_void_p, c_int, c_int, c_void_p, c_size_t]),
    "rt_debug_set_trel": (c_int, [c_void_p, P(c_float)]),
    "rt_debug_prepare": (c_int, [c_void_p]),
    "rt_debug_linearize": (c_int, [c_void_p, c_int, c_int]),
    "rt_debug_depth_reciprocals": (c_int, [c_int, P(c_float), c_int]),
    "rt_last_error": (ctypes.c_char_p, []),
    "rt_cnn_default_config": (None, [P(RtCnnConfig)]),
    "rt_cnn_create": (c_int, [c_int, P(RtCnnConfig), c_int, P(c_void_p)]),
    "rt_cnn_destroy": (None, [c_void_p]),
    "rt_cnn_num_params": (c_size_t, [c_void_p]),
    "rt_cnn_set_params": (c_int, [c_void_p, P(c_float)]),
    "rt_cnn_get_params": (c_int, [c_void_p, P(c_float)]),
    "rt_cnn_stream": (c_void_p, [c_void_p]),
    "rt_cnn_make_inputs": (c_int, [c_void_p, c_int, P(RtPair), c_void_p]),
    "rt_cnn_forward": (c_int, [c_void_p, c_int, c_void_p, c_void_p]),
    "rt_cnn_loss_grad": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, P(c_float), P(c_float), P(c_float)]),
    "rt_cnn_train_step": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_float, P(c_float)]),
    "rt_cnn_augment": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, P(c_int), c_void_p, c_void_p]),
    "rt_cnn_pose_to_target": (c_int, [P(RtCnnConfig), c_int, P(c_double), P(c_float)]),
    "rt_cnn_target_to_pose": (c_int, [P(RtCnnConfig), c_int, P(c_float), P(c_double)]),
    "rt_batch_prepare": (c_int, [c_void_p, c_int, P(RtPair)]),
    "rt_batch_propose": (c_int, [c_void_p, c_void_p]),
    # include/rt_sequence.h
    "rt_sequence_default_params": (None, [c_void_p]),
    "rt_sequence_create": (c_int, [c_int, c_int, c_int, P(c_float), c_void_p, c_void_p, P(c_void_p)]),
    "rt_sequence_destroy": (None, [c_void_p]),
    "rt_sequence_run": (c_int, [c_void_p, c_void_p, c_void_p, c_int, P(c_double), P(c_int), P(c_float), P(c_int)]),
    # include/rt_sensor.h
    "rt_calib_load_xml": (c_int, [ctypes.c_char_p, c_int, c_int, P(c_float), P(c_int)]),
    "rt_calib_parse_xml": (c_int, [ctypes.c_char_p, c_size_t, c_int, c_int, P(c_float), P(c_int)]),
    "rt_pnm_info": (c_int, [ctypes.c_char_p, P(c_int), P(c_int), P(c_int), P(c_int)]),
    "rt_pnm_read": (c_int, [ctypes.c_char_p, c_void_p, c_size_t]),
    "rt_disparity_to_depth": (c_int, [c_void_p, c_int, c_int, P(c_float), c_int, c_int, c_int, c_void_p, c_void_p]),
    "rt_depth_mm_to_u16": (c_int, [c_void_p, c_size_t, c_void_p, c_void_p]),
    # include/rt_trajectory.h
    "rt_trajectory_write": (c_int, [ctypes.c_char_p, c_int, P(c_float), c_int, c_int]),
    "rt_trajectory_read": (c_int, [ctypes.c_char_p, c_int, c_int, c_int, P(c_float), P(c_int)]),
    "rt_schedule_read": (c_int, [ctypes.c_char_p, c_int, c_void_p, P(c_int)]),
    "rt_schedule_lookup": (c_int, [c_void_p, c_int, c_int, P(c_int), P(c_float)]),
    "rt_pose_distance": (c_int, [P(c_float), P(c_double), P(c_double)]),
    "rt_grid_object": (c_int, [c_int, c_int, c_int, c_float, c_float, c_float, c_float, P(c_float)]),
    "rt_keyframes_similar": (c_int, [P(c_float), c_int, P(c_float), c_float, c_float, P(c_float), P(c_float), c_int,
                                     P(c_int), P(c_float), P(c_int)]),
    # include/rt_tsdf.h
    "rt_tsdf_default_params": (None, [c_void_p]),
    "rt_tsdf_create": (c_int, [c_int, c_void_p, P(c_void_p)]),
    "rt_tsdf_destroy": (None, [c_void_p]),
    "rt_tsdf_reset": (c_int, [c_void_p]),
    "rt_tsdf_set_depth_rgbd": (c_int, [c_void_p, c_void_p]),
    "rt_tsdf_set_depth_u16": (c_int, [c_void_p, c_void_p]),
    "rt_tsdf_update": (c_int, [c_void_p, P(c_float), P(c_float)]),
    "rt_tsdf_raycast": (c_int, [c_void_p, P(c_float), P(c_float), c_int, c_void_p, c_void_p, c_void_p]),
    "rt_tsdf_resample": (c_int, [c_void_p, c_void_p, P(c_float)]),
    "rt_tsdf_read": (c_int, [c_void_p, c_void_p, c_size_t]),
    "rt_tsdf_data": (c_void_p, [c_void_p]),
    "rt_tsdf_stream": (c_void_p, [c_void_p]),
    "rt_tsdf_synchronize": (c_int, [c_void_p]),
}

GEOMETRIC, PHOTOMETRIC, BIOBJECTIVE = 0, 1, 2
ROBUST_TUKEY, ROBUST_HUBER = 0, 1
SOLVER_CHOLESKY, SOLVER_CG = 0, 1


def lib():
    """The HIP tracker library; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m recon_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            if not hasattr(L, name):
                continue
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def synth_lib():
    """The synthetic-scene entry points (rs_*) of the same library."""
    global _synth
    if _synth is None:
        S = lib()
        S.rs_scene_create.restype = c_void_p
        S.rs_scene_create.argtypes = [ctypes.c_uint64]
        S.rs_scene_destroy.argtypes = [c_void_p]
        S.rs_render.argtypes = [c_void_p, P(c_double), c_int, c_int, c_double, c_double, c_double, c_double,
                                ctypes.c_uint64, c_double, c_void_p, c_void_p]
        S.rs_sample_motion.argtypes = [ctypes.c_uint64, c_double, c_double, P(c_double)]
        S.rs_trajectory.argtypes = [ctypes.c_uint64, c_int, c_double, c_double, c_int, c_double, P(c_double)]
        S.rs_render_device.restype = c_int
        S.rs_render_device.argtypes = [c_void_p, P(c_double), c_int, c_int, c_double, c_double, c_double, c_double,
                                       ctypes.c_uint64, c_double, c_void_p, c_void_p, c_void_p]
        _synth = S
    return _synth


def check(rc, what="call"):
    if rc is not None and rc < 0:
        msg = lib().rt_last_error()
        raise RuntimeError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
    return rc
