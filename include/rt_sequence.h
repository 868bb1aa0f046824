/*
 * rt_sequence.h -- keyframe tracking of a frame sequence with learned warm
 * starts, the plain tracker as fallback and arbiter, and batched
 * relocalisation against recent keyframes.
 *
 * The reference drives its tracker per frame from the application
 * (rgbdscan/src/main.cpp:923-986: set the current image, track, take a new
 * keyframe when PhotoTracker::referenceOutdated, phototracker.cpp:292-296)
 * and searches its keyframe database by pose distance
 * (KeyFrameModel::findSimilarKeyFrames, reconstruct/KeyFrameModel.cpp:1476-1490).
 * This driver composes the same steps over the library (include/rt_tracker.h,
 * rt_cnn.h, rt_trajectory.h) for C and C++ callers:
 *
 * per frame k (1 .. n-1):
 *   1. with a network, its proposal of the camera pose relative to the
 *      keyframe starts Gauss-Newton (rt_track_from); without, the previous
 *      estimate does (rt_track);
 *   2. when the frame's weighted residual RMS exceeds mean + reloc_sigma *
 *      std of the last reloc_window frames (after 10 frames), the frame is
 *      also tracked from the previous frame's pose and the lower RMS wins;
 *   3. if it still spikes, the frame is tracked against up to
 *      reloc_keyframes keyframes in one batched launch from the network's
 *      proposals (first those the keyframe search finds within
 *      reloc_search_mm / reloc_search_deg of the current estimate, most
 *      similar first, then the most recent others); the reloc_verify best are
 *      re-tracked by a second single-pair tracker and tracking moves to the
 *      best one that fits better and agrees with the current estimate within
 *      reloc_gate_deg / reloc_gate_mm;
 *   4. the fuse_frames frames after a keyframe are fused into it, and the
 *      keyframe refreshed after the last of them; the current frame becomes
 *      the keyframe when its pose relative to the keyframe exceeds dist_mm /
 *      angle_deg or after max_interval frames.
 * Returns RT_OK (0) or a negative rt_status; rt_last_error() describes it.
 */
#ifndef RT_SEQUENCE_H
#define RT_SEQUENCE_H

#include <stdint.h>

#include "rt_cnn.h"
#include "rt_tracker.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int mode;                 /* RT_BIOBJECTIVE */
    float lambda;             /* 0.5 */
    int n_select;             /* 61440 photometric points of the single-pair trackers */
    int fuse_frames;          /* 5 */
    int max_interval;         /* 10 */
    float dist_mm, angle_deg; /* 60 mm, 5 deg: a new keyframe beyond these */
    int relocalise;           /* 1 (needs a network) */
    int reloc_keyframes;      /* 64 keyframes per batched relocalisation */
    float reloc_sigma;        /* 3 */
    int reloc_window;         /* 30 frames of RMS statistics */
    int reloc_points;         /* 8192 points per relocalisation pair (photometric) */
    int reloc_verify;         /* 3 candidates re-tracked */
    float reloc_gate_deg, reloc_gate_mm;     /* 20 deg, 300 mm */
    float reloc_search_mm, reloc_search_deg; /* 1000 mm, 45 deg */
    int reloc_history;        /* 1024 keyframes remembered */
} rt_sequence_params;

typedef struct rt_sequence rt_sequence;

void rt_sequence_default_params(rt_sequence_params* p);
/* net: the pose network (rt_cnn.h, max_batch >= reloc_keyframes), or NULL for
 * the plain tracker (previous-estimate starts, no relocalisation). The
 * network is borrowed and must outlive the sequence tracker. */
int rt_sequence_create(int device, int width, int height, const float calib[RT_CALIB_SIZE], rt_cnn* net,
                       const rt_sequence_params* params, rt_sequence** out);
void rt_sequence_destroy(rt_sequence* s);
/* Track n frames held in device memory: frame k is (gray + k*W*H, depth +
 * k*W*H), 8-bit gray and 16-bit depth in mm. Frame 0 is the first keyframe at
 * the identity. poses_out (host, n x 16 doubles, row-major) receives the
 * global camera poses; trace_out (host, nullable, n x 2 ints) per frame its
 * keyframe and how it was tracked (0 first estimate, 1 plain fallback, 2
 * relocalised); rms_out (host, nullable, n floats) its residual RMS.
 * stats (nullable, 2 ints): keyframes taken, relocalisations. */
int rt_sequence_run(rt_sequence* s, const uint8_t* gray_dev, const uint16_t* depth_dev, int n, double* poses_out,
                    int* trace_out, float* rms_out, int* stats);

#ifdef __cplusplus
}
#endif
#endif /* RT_SEQUENCE_H */
