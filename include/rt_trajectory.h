/*
 * rt_trajectory.h -- the files and the keyframe search around the tracker:
 * the trajectory file the reference writes (cameraMatrixEst.txt), its
 * per-frame tracking schedule (best-method.tst), and the pose-distance search
 * of similar keyframes that relocalisation draws its candidates from.
 *
 * Poses are row-major 4x4 floats. Every entry point returns RT_OK (0) or a
 * negative rt_status; rt_last_error() gives the message.
 */
#ifndef RT_TRAJECTORY_H
#define RT_TRAJECTORY_H

#include <stddef.h>
#include <stdint.h>

#include "rt_tracker.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Trajectory file. Replaces ResultTxt::canonize + ResultTxt::save
 * (tracker/calib/ResultTxt.cpp:52-76): every pose is left-multiplied by the
 * inverse of the first (invertRT4, matrixMult4x4 in float), then optionally
 * inverted (`inverse`) and transposed (`transpose`), and written as one line
 * "k: m0 m4 m8 m12 m1 m5 m9 m13 m2 m6 m10 m14 m3 m7 m11 m15" (printf %e, k from
 * 1) -- the matrix column by column. n >= 1 poses. */
int rt_trajectory_write(const char* path, int n, const float* poses, int transpose, int inverse);
/* Read such a file back: the written matrices (after canonisation and the
 * same `transpose` / `inverse` undone), up to max poses; *n receives the
 * count. Values carry the file's 7 significant digits. */
int rt_trajectory_read(const char* path, int transpose, int inverse, int max, float* poses, int* n);

/* One line of a tracking schedule: from `frame` on, track in `mode`
 * (RT_GEOMETRIC, RT_PHOTOMETRIC, RT_BIOBJECTIVE; -1 for an unknown mode name)
 * with `lambda`; `error` is the third column (a residual the schedule was
 * chosen by). */
typedef struct {
    int frame;
    int mode;
    float lambda;
    float error;
} rt_schedule_entry;
/* Read a schedule file. Replaces TrackingParams::init
 * (tracker/calib/TrackingParams.cpp:50-94): lines "frame mode error" with mode
 * "geometric" (lambda 1), "photometric" (lambda 0) or
 * "biobjective-lambdaX" (lambda X). Up to max entries; *n receives the count. */
int rt_schedule_read(const char* path, int max, rt_schedule_entry* out, int* n);
/* Mode and lambda for a frame: the entry whose frame is nearest (the first of
 * equally near ones). Replaces TrackingParams::getParams
 * (TrackingParams.cpp:96-110); with n = 0 the reference's default,
 * RT_GEOMETRIC with lambda 1. */
int rt_schedule_lookup(const rt_schedule_entry* entries, int n, int frame, int* mode, float* lambda);

/* Distance and angle of a relative pose dT. Replaces poseDistance
 * (tracker/tracker/basic_math.cpp:1070-1094): dist = |t| (+1e-12 under the
 * root, rounded to float), angle = the largest absolute Euler angle of
 * rot2euler (basic_math.cpp:1059-1068), in degrees; FP64. */
int rt_pose_distance(const float* dT, double* dist, double* angle_deg);

/* Test object of the keyframe search: nx*ny*nz points, layer k at depth
 * -view_min - k*z_range/(nz-1) (one layer: at -view_min - z_range/2), a
 * 2*size square grid scaled with depth. Replaces
 * KeyFrameModel::generateGridObject (reconstruct/KeyFrameModel.cpp:707-734);
 * out holds 3*nx*ny*nz floats. */
int rt_grid_object(int nx, int ny, int nz, float size, float view_min, float z_range, float fov_x_deg, float* out);

/* Similar keyframes of a camera pose T among nkeys keyframe poses: those
 * whose relative pose inv(T) * key passes rt_pose_distance within
 * (max_dist, max_angle), ordered by the mean squared image displacement of the
 * test object under the relative pose, projected with K (3x3, the colour
 * intrinsics of the calibration block), smallest first; weights[i] = 1 - (w_i
 * - w_min) / (w_max - w_min) (NaN when all are equal, as in the reference).
 * Replaces KeyFrameModel::enumerateNearPoses / findSimilarKeyFrames /
 * projectionDistance / bubbleSort (KeyFrameModel.cpp:1423-1451, 1476-1490,
 * 1530-1600). order and weights hold nkeys entries; *count receives the
 * number found. */
int rt_keyframes_similar(const float* T, int nkeys, const float* key_poses, float max_dist, float max_angle,
                         const float* K, const float* grid, int nvertices, int* order, float* weights, int* count);

#ifdef __cplusplus
}
#endif

#endif /* RT_TRAJECTORY_H */
