/*
 * rt_tracker.h -- C ABI of the MI355X-native RGB-D keyframe tracker.
 *
 * Drop-in boundary for the pose-estimation path of ttykkala/recon: the
 * PhotoTracker host API (tracker/tracker/phototracker.h:33-82) with its
 * ShapeTracker behind it (tracker/tracker/shapetracker.h:29-72). All device work
 * runs as hand-written HIP kernels for gfx950 on one stream; the tracker owns
 * its device memory. Every function returns RT_OK (0) or a negative rt_status
 * unless stated otherwise; rt_last_error() describes the last failure of the
 * calling thread. A tracker handle is used from one thread at a time; there is
 * no hidden thread and no barrier (the reference runs its tracker on a pthread
 * synchronised by a 2-party barrier, phototracker.cpp:691-728, and
 * barrier_sync.h:27-70; here the caller's thread drives it directly).
 *
 * Conventions are the reference's: camera frame x right, y up, looking down -z;
 * depth and translations in millimetres; poses row-major 4x4. The calibration
 * is the reference's 96-float block (tracker/calib/calib.h:20-36), intrinsics
 * already multiplied by the image width (calib.cpp:74-93), fx < 0.
 */
#ifndef RT_TRACKER_H
#define RT_TRACKER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* calibration block layout (tracker/calib/calib.h:20-36) */
#define RT_CALIB_SIZE 96
#define RT_KR_OFFSET 0
#define RT_IKR_OFFSET 9
#define RT_KCR_OFFSET 18
#define RT_TLR_OFFSET 26
#define RT_C1_OFFSET 42
#define RT_KL_OFFSET 43
#define RT_IKL_OFFSET 52
#define RT_C0_OFFSET 61
#define RT_MIND_OFFSET 62
#define RT_MAXD_OFFSET 63
#define RT_TRL_OFFSET 64

/* TrackMode (tracker/tracker/phototracker.h:31), same order */
enum { RT_GEOMETRIC = 0, RT_PHOTOMETRIC = 1, RT_BIOBJECTIVE = 2 };
/* photometric M-estimator: Tukey as in histogram64.cu:79-94, or Huber */
enum { RT_ROBUST_TUKEY = 0, RT_ROBUST_HUBER = 1 };
/* 6x6 solver: FP64 Cholesky, or the reference's 6-step FP64 CG (warp.cu:926-975) */
enum { RT_SOLVER_CHOLESKY = 0, RT_SOLVER_CG = 1 };

typedef enum {
    RT_OK = 0,
    RT_ERR_ARG = -1,      /* invalid argument (null, size, range) */
    RT_ERR_HIP = -2,      /* a HIP runtime call failed */
    RT_ERR_STATE = -3,    /* call not valid in the current state */
    RT_ERR_NOMEM = -4,    /* device or host allocation failed */
    RT_ERR_UNSUPPORTED = -5
} rt_status;

/* Tracker parameters. rt_default_params() fills the reference's values:
 * mode GEOMETRIC and lambda 0.9 (phototracker.cpp:134-135), levels from the
 * width (4 at 640, else 3; phototracker.cpp:777-784), iterations {1,3,10,5}
 * coarse to fine (:148-149), 8192 photometric points (:48), 64-bin MAD Tukey
 * (histogram64.cu), keyframe thresholds 125 mm / 12.5 deg (:127-128). */
typedef struct {
    int mode;
    float lambda;
    int levels;          /* 0: width rule */
    int iters[4];        /* coarse -> fine; iters[levels-1-l] is used at level l */
    int n_select;        /* photometric points, a multiple of 256, <= W*H/2 */
    int mad_bins;        /* 64 (reference) or 256 */
    int robust;          /* RT_ROBUST_* */
    int solver;          /* RT_SOLVER_* */
    float pose_dist_threshold;   /* mm */
    float pose_angle_threshold;  /* degrees */
    int incremental;             /* setIncrementalMode */
    int reference_updating;      /* setReferenceUpdating */
} rt_params;

typedef struct rt_tracker rt_tracker;

void rt_default_params(rt_params* p);

/* A pinhole calibration block in the reference convention for intrinsics in
 * pixels with fx, fy > 0 and image rows top-down: KL = KR = [[-fx,0,cx],[0,fy,cy],
 * [0,0,1]], no distortion, TLR = TRL = identity, depth window [zmin, zmax] mm. */
int rt_calib_pinhole(float fx, float fy, float cx, float cy, float zmin, float zmax, float out[RT_CALIB_SIZE]);

/* initializePhotoTracker + allocMemory (phototracker.cpp:900-913, 757-847).
 * width and height must be multiples of 2^(levels-1) and width*height/512 must
 * be at most 1024 (the selection segments, filterIndex256.cu:110-169). */
int rt_create(int device, int width, int height, const float calib[RT_CALIB_SIZE], const rt_params* params,
              rt_tracker** out);
void rt_destroy(rt_tracker* t);

/* setCurrentImage (phototracker.cpp:256-266). gray: 8-bit, depth: 16-bit mm
 * (0 = no measurement), both width*height, rows top-down. on_device != 0: the
 * pointers are device memory of the tracker's device (no host copy). */
int rt_set_current(rt_tracker* t, const uint8_t* gray, const uint16_t* depth, int on_device);
/* The reference's own input: a float4 RGBD image (rgb in [0,1], w = depth
 * normalised to the calib window) stored bottom-up as an OpenGL texture and read
 * Y-flipped (rgb2gray.cu:64-76). */
int rt_set_current_rgbd(rt_tracker* t, const float* rgbd, int on_device);

/* setReferencePoints with raw measurements (phototracker.cpp:310-337,
 * shapetracker.cpp:195-211). Updates the keyframe when force != 0 or the
 * reference is outdated for the suggested mode (referenceOutdated, :292-296);
 * returns 1 when the keyframe was replaced, 0 when not, < 0 on error. */
int rt_set_reference(rt_tracker* t, const uint8_t* gray, const uint16_t* depth, int on_device, int suggested_mode,
                     float lambda, int force);
int rt_set_reference_rgbd(rt_tracker* t, const float* rgbd, int on_device, int suggested_mode, float lambda,
                          int force);

/* PhotoTracker::setFilteredReference (phototracker.cpp:71-74,
 * shapetracker.cpp:59-65): with the flag on, the ICP keyframe comes from
 * ray-cast points and normals (the TSDF model, include/rt_tsdf.h) instead of
 * the measured depth; the photometric keyframe is the frame either way. Off
 * by default, as in the reference application (rgbdscan/src/main.cpp:1236).
 * While on, rt_set_reference / rt_set_reference_rgbd fail (the reference's
 * setReferencePoints returns false without point maps, phototracker.cpp:313)
 * and so does rt_track_sequence, whose keyframes are measured frames. */
int rt_set_filtered_reference(rt_tracker* t, int flag);
/* Inspection images for display. updateErrorImage (phototracker.cpp:388-403):
 * photometric mode writes min(25 |r|, 1) of each selected point's level-0
 * residual under the current pose at its pixel (other pixels keep their
 * values; xyz2DiffKernel, warp.cu:1955-1999); geometric and bi-objective
 * modes write, per level, min(|z' - z_cur| / 30, 1) against the current point
 * the keyframe point projects to, 0 elsewhere (xyz6DiffNormalizedKernel,
 * rgb2gray.cu:143-171). The keyframe must be prepared (after rt_track or
 * rt_debug_prepare). */
int rt_update_error_image(rt_tracker* t);
/* getDebugImage1C (:734-744) and getRefImage1C (:746-755), copied to host
 * memory of `count` floats: photometric mode the W x H error image and the
 * keyframe's depth clamped to [min, max] and mapped to [0, 1] (xyz2Depth-
 * NormalizedKernel, rgb2gray.cu:104-116); the other modes the level's
 * geometric error image and the shape tracker's reference image, which the
 * reference never writes (zeros). */
int rt_get_debug_image(rt_tracker* t, int level, float* out, size_t count);
int rt_get_ref_image(rt_tracker* t, int level, float* out, size_t count);
int rt_get_filtered_reference(const rt_tracker* t);  /* getFilteredReference :82-84 */
/* PhotoTracker::setReferencePoints(xyz, normals, frame, mode, lambda)
 * (phototracker.cpp:310-337) with the ray-cast maps: xyz and normals are
 * device float4 images of width x height (camera coordinates of the
 * keyframe; xyz.w and normals.w are not read), e.g. rt_tsdf_raycast's outputs.
 * With the filtered reference on, the ICP cloud keeps the points with z < 0
 * and |n|^2 > 1e-7 (generateOrientedPoints6Cuda, d2z.cu:471-502,
 * 1169-1181); with it off, the maps are ignored and the call is
 * rt_set_reference / rt_set_reference_rgbd. Return values as there. */
int rt_set_reference_points(rt_tracker* t, const float* xyz, const float* normals, const uint8_t* gray,
                            const uint16_t* depth, int on_device, int suggested_mode, float lambda, int force);
int rt_set_reference_points_rgbd(rt_tracker* t, const float* xyz, const float* normals, const float* rgbd,
                                 int on_device, int suggested_mode, float lambda, int force);

/* readyToTrack -> gpuPreProcess -> gpuEstimateIncrement -> applyIncrement
 * (phototracker.cpp:704-718). Returns 1 when a frame was tracked, 0 when not
 * ready (no keyframe or no current frame for the mode), < 0 on error. */
int rt_track(rt_tracker* t);

/* rt_track with the Gauss-Newton loop started from T_init, the current camera
 * pose relative to the keyframe (row-major 4x4, mm), instead of the previous
 * estimate. No counterpart in the reference: the warm start of a learned pose
 * proposal. Returns as rt_track. */
int rt_track_from(rt_tracker* t, const double T_init[16]);
/* Weighted photometric residual RMS of the last iteration, sqrt(sum (w r)^2 /
 * sum w^2) over the selected points (photometric and bi-objective modes). No
 * counterpart in the reference: the trigger of relocalisation. */
int rt_get_residual_rms(rt_tracker* t, float* rms);

/* getGlobalPose / getRelativePose / getIncrement (phototracker.cpp:223-229, 564) */
int rt_get_global_pose(const rt_tracker* t, double T[16]);
int rt_get_relative_pose(const rt_tracker* t, float T[16]);
int rt_get_increment(const rt_tracker* t, double T[16]);

int rt_reset_tracking(rt_tracker* t, int mode);                  /* resetTracking :231-243 */
int rt_clear_pose(rt_tracker* t);                                /* clearPose :245-253 */
int rt_reset_to_transform(rt_tracker* t, const double T[16]);    /* resetToTransform :545-561 */
int rt_set_mode(rt_tracker* t, int mode);                        /* setMode :174-181 */
int rt_get_mode(const rt_tracker* t);
int rt_set_lambda(rt_tracker* t, float lambda);                  /* setLambda :163-168, clamped to [0,1] */
float rt_get_lambda(const rt_tracker* t);
int rt_set_iteration_counts(rt_tracker* t, const int n[4]);      /* setIterationCounts :494-496 */
int rt_get_iteration_counts(const rt_tracker* t, int n[4]);      /* getIterationCounts :638-644 */
int rt_set_reference_updating(rt_tracker* t, int flag);          /* setReferenceUpdating :57-64 */
int rt_set_incremental_mode(rt_tracker* t, int flag);            /* setIncrementalMode :71-75 */
int rt_signal_calibration_update(rt_tracker* t, const float calib[RT_CALIB_SIZE]); /* :571-574 */
int rt_get_layers(const rt_tracker* t);                          /* getLayers :489-491 */
int rt_num_reference_updates(const rt_tracker* t);               /* numReferenceUpdates :188-190 */

/* listSelected (phototracker.cpp:405-446): image coordinates of the selected
 * photometric points projected with the identity (ref_xy) and with the current
 * relative pose (cur_xy), and their robust weights of the last iteration.
 * capacity: entries available in each output; *n receives the count. */
int rt_list_selected(rt_tracker* t, float* cur_xy, float* ref_xy, float* weight, int capacity, int* n);

/* Keyframe fusion: a running weighted mean of the current frame's depth and
 * gray into the keyframe, through the current relative pose (per pixel, 30 mm
 * gate, weight cap 16). rt_refresh_keyframe re-prepares the keyframe from the
 * fused images, keeping the relative pose. */
int rt_fuse_current(rt_tracker* t);
int rt_refresh_keyframe(rt_tracker* t);

/* The tracker's HIP stream (hipStream_t) and a full synchronisation. */
void* rt_stream(rt_tracker* t);
int rt_synchronize(rt_tracker* t);

/* Sequence driver: track n frames held in device memory back to back with no
 * host synchronisation per frame. Frame k is (gray + k*W*H, depth + k*W*H).
 * Frame 0 becomes the keyframe; every keyframe_interval frames the previous
 * frame becomes the new keyframe; the fuse_frames frames after a keyframe are
 * fused into it and the keyframe is refreshed after the last of them. Global
 * poses (double[16] per frame, frame 0 = identity) are written to poses_out
 * (host memory) when the call returns. */
int rt_track_sequence(rt_tracker* t, const uint8_t* gray_dev, const uint16_t* depth_dev, int n, int keyframe_interval,
                      int fuse_frames, double* poses_out);

/* ------------------------------------------------------------------ batches
 * Many independent frame pairs of mixed sizes tracked by one persistent
 * launch: each workgroup owns one pair at a time and runs its whole
 * coarse-to-fine Gauss-Newton loop, taking pairs from a device work queue, in
 * the mode of the batch's parameters: photometric, geometric (point-to-plane
 * ICP with Student-t weights) or bi-objective (both systems fused by
 * linearFuse with params->lambda), as ShapeTracker / PhotoTracker
 * (shapetracker.cpp:357-394, phototracker.cpp:666-689) do for one pair. No
 * reference counterpart for the batching: the reference tracks one pair. */
typedef struct {
    const uint8_t* ref_gray;    /* device pointers, width*height each */
    const uint16_t* ref_depth;
    const uint8_t* cur_gray;
    const uint16_t* cur_depth;
    int width, height;
    const float* calib;         /* host pointer to a 96-float block */
    const double* init_T;      /* host, nullable: initial ref->cur point transform (Trel) */
} rt_pair;

typedef struct {
    double Trel[16];   /* converged ref->cur point transform */
    double T[16];      /* its inverse: the current camera in the reference frame */
    float rms;         /* weighted residual RMS of the last iteration */
    int status;        /* 0 ok, < 0 degenerate system */
} rt_pair_result;

typedef struct rt_batch rt_batch;

int rt_batch_create(int device, int max_pairs, int max_width, int max_height, const rt_params* params,
                    rt_batch** out);
void rt_batch_destroy(rt_batch* b);
/* Prepare keyframes and current pyramids for n pairs and run the tracking
 * launch; results are copied to out (host) when the call returns. */
int rt_batch_track(rt_batch* b, int n, const rt_pair* pairs, rt_pair_result* out);
/* Re-run only the tracking launch on the pairs of the last rt_batch_track /
 * rt_batch_prepare (inputs already prepared in device memory), from the given
 * initial poses (n x 16 doubles, host), or, with init_T NULL, from the initial
 * poses already on the device (those of the last call, or the proposals of
 * rt_batch_propose). out NULL: enqueue without waiting. */
int rt_batch_retrack(rt_batch* b, const double* init_T, rt_pair_result* out);
/* Wait for the batch's stream and copy the results of its last tracking
 * launch (one that rt_batch_retrack enqueued with out NULL) to out (host).
 * No counterpart in the reference: it lets a caller keep two batches in
 * flight, preparing one while the other tracks. */
int rt_batch_results(rt_batch* b, rt_pair_result* out);
void* rt_batch_stream(rt_batch* b);
/* The persistent kernel's variant table (the autotuner's candidates; the
 * environment variable RT_BATCH_VARIANT=v pins variant v): info receives
 * {workgroup size, points per thread (unroll depth where residuals live in
 * LDS), residuals in LDS (0/1), LDS point-cache capacity in points, waves per
 * SIMD, tracking mode}. Returns RT_ERR_ARG past the end of the table. */
int rt_batch_variant_info(int v, int info[6]);
/* Inspection of one pair of the last prepared batch (for parity tests):
 * RT_DBG_GRAD_HIST, RT_DBG_SELECTED, RT_DBG_POINTS, RT_DBG_JACOBIAN (level),
 * RT_DBG_TREL (the pair's initial pose), RT_DBG_DEPTH_BLOCKS; in the geometric and bi-objective
 * modes also RT_DBG_REF_CLOUD, RT_DBG_CUR_CLOUD (level) and, after tracking,
 * RT_DBG_ICP_COUNT and RT_DBG_ICP_VAR (level) of the last tracking launch;
 * layouts as for single trackers. */
int rt_batch_debug_read(rt_batch* b, int pair, int what, int level, void* host, size_t bytes);

/* Inspection for bounds checks on the host (no device is touched): the pair
 * descriptors rt_batch_prepare builds for n pairs in a batch of max_width x
 * max_height, as byte offsets from the batch's arena. fields: 40 int64 per
 * pair: W, H, levels, points, simple-calibration flag, offsets of the gray
 * pyramid levels 1-3 of the keyframe and of the current frame (-1 where the
 * level is absent), of the gradient bins, gradient histogram, selection,
 * points and the point records of levels 0-3, the pair's arena slot start
 * and the slot size, then in the geometric modes (else -1) the offsets of the
 * keyframe cloud, its normals and the current cloud of levels 0-3, of the ICP
 * selection, its squared residuals, the ICP counts and variances, the
 * offset of the current depth map in 8x8-pixel blocks, then 2 zeros. order:
 * the work order (n ints); maxes: the largest width, height and point count. */
int rt_batch_layout(const rt_params* params, int max_width, int max_height, int n, const rt_pair* pairs,
                    int64_t* fields, int* order, int maxes[3]);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): rt_batch_track without a batch object. The kernels'
 * working memory (per-pair pyramids, gradient bins, histograms, selections,
 * point records, descriptor and pose tables, the work queue) is one fixed
 * device array of RT_NOALLOC_SCRATCH_BYTES (21.6 MiB) that each GPU the library
 * runs on holds until the process ends; the call allocates nothing. The caller
 * passes the frames (device pointers), calibration and initial poses as in
 * rt_pair, and the results; work runs on `stream` (a hipStream_t, NULL for the
 * null stream) and the call returns when the results are in `out`. Limits:
 * n <= RT_NOALLOC_MAX_PAIRS, pairs up to RT_NOALLOC_MAX_WIDTH x
 * RT_NOALLOC_MAX_HEIGHT, params->n_select <= RT_NOALLOC_MAX_POINTS (RT_ERR_ARG
 * beyond). Every call on one GPU uses that GPU's one array: calls that overlap
 * on one GPU (other threads, other streams) race. The tracking kernel variant
 * is fixed (no autotuning); poses agree with rt_batch_track's within rounding. */
#define RT_NOALLOC_MAX_PAIRS 8
#define RT_NOALLOC_MAX_WIDTH 640
#define RT_NOALLOC_MAX_HEIGHT 480
#define RT_NOALLOC_MAX_POINTS 8192
#define RT_NOALLOC_SCRATCH_BYTES 22598656
int rt_batch_track_noalloc(int device, int n, const rt_pair* pairs, const rt_params* params, void* stream,
                           rt_pair_result* out);

/* ------------------------------------------------------------------ inspection
 * Read an internal device buffer into host memory, for parity tests.
 * what: RT_DBG_*; level: pyramid level where applicable. */
enum {
    RT_DBG_REF_GRAY = 0,     /* float W_l*H_l */
    RT_DBG_CUR_GRAY = 1,     /* float W_l*H_l */
    RT_DBG_REF_Z = 2,        /* float W*H */
    RT_DBG_CUR_Z = 3,        /* float W*H */
    RT_DBG_GRAD_MAG = 4,     /* float W*H (vertex slot 13) */
    RT_DBG_GRAD_HIST = 5,    /* uint32 256 */
    RT_DBG_SELECTED = 6,     /* int32 n_select */
    RT_DBG_POINTS = 7,       /* float4 n_select: x, y, z, 0 */
    RT_DBG_JACOBIAN = 8,     /* per level, n_select points: float4 {J0..J3}, then float2 {J4,J5}, then float I_ref */
    RT_DBG_RESIDUAL = 9,     /* float2 n_select: r, zWeight (last photometric iteration) */
    RT_DBG_TREL = 10,        /* float 16 */
    RT_DBG_REF_CLOUD = 11,   /* float4 2*W_l*H_l: (x,y,z,mask),(nx,ny,nz,0) per pixel */
    RT_DBG_CUR_CLOUD = 12,   /* float4 W_l*H_l: (x,y,z,mask) */
    RT_DBG_ICP_SELECTED = 13,/* int32 W_l*H_l (count in RT_DBG_ICP_COUNT) */
    RT_DBG_ICP_COUNT = 14,   /* int32 4 (per level) */
    RT_DBG_ICP_R2 = 15,      /* float W_l*H_l: squared residuals of the compacted points */
    RT_DBG_ICP_VAR = 16,     /* float 3 per level: variance sums of the 3 Student-t rounds */
    RT_DBG_SYSTEM = 17,      /* double 54: photometric terms then geometric terms, 27 each:
                                 A upper triangle row by row (21), then b (6) */
    RT_DBG_KF_WEIGHT = 18,   /* float W*H fusion weights */
    RT_DBG_MAD_HIST = 19,    /* uint32 mad_bins (last photometric iteration) */
    RT_DBG_DEPTH_BLOCKS = 20 /* batches: uint16 W*ceil8(H), the current depth map in 8x8-pixel blocks:
                                pixel (x, y) at ((y/8)*(W/8) + x/8)*64 + (y%8)*8 + x%8 */
};
int rt_debug_read(rt_tracker* t, int what, int level, void* host, size_t bytes);
/* Set the device relative pose (TrelDev) directly. */
int rt_debug_set_trel(rt_tracker* t, const float T[16]);
/* Prepare the keyframe and the current frame (gpuPreProcess) without tracking. */
int rt_debug_prepare(rt_tracker* t);
/* One linearisation at the current Trel, no update: photometric system of
 * level l and/or the geometric one (iteration it: 0 re-selects and re-weights).
 * Results are readable through RT_DBG_SYSTEM, RT_DBG_RESIDUAL, RT_DBG_ICP_*. */
int rt_debug_linearize(rt_tracker* t, int level, int it);
/* The reciprocal the batched preparation takes of each 16-bit depth value:
 * host[i] = RN(1 / -(i + 1)) as the device computes it, for i < n <= 65535.
 * No counterpart in the reference (whose kernels divide); it lets a test
 * compare the device reciprocal with IEEE division over every input. */
int rt_debug_depth_reciprocals(int device, float* host, int n);

const char* rt_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* RT_TRACKER_H */
