/*
 * rt_tsdf.h -- truncated signed distance fusion of depth frames into a voxel
 * grid and ray casting of the grid into point / normal maps: the keyframe
 * model the reference's application keeps beside the tracker
 * (reconstruct/VoxelGrid.cpp, kernels/VoxelGrid_cuda.cu), whose ray-cast maps
 * it hands to PhotoTracker::setReferencePoints as the filtered reference.
 *
 * Poses are row-major 4x4 floats, camera -> world (the tracker's global
 * pose); K is the depth camera's intrinsics normalised by the image size
 * (K[0] = fx / width, K[2] = cx / width, K[4] = fy / height, K[5] = cy /
 * height, as getNormalizedK builds it, rgbdscan/src/main.cpp:878-882).
 * Device buffers are plain device pointers; every entry point returns RT_OK
 * (0) or a negative rt_status.
 *
 * Work is queued on the grid's own stream (rt_tsdf_stream) and the calls
 * return without waiting, except rt_tsdf_create, rt_tsdf_reset and
 * rt_tsdf_read: device inputs must stay valid, and outputs may be read, only
 * after rt_tsdf_synchronize (or other work ordered after the stream).
 */
#ifndef RT_TSDF_H
#define RT_TSDF_H

#include <stddef.h>
#include <stdint.h>

#include "rt_tracker.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rt_tsdf rt_tsdf;

typedef struct {
    int resolution;      /* voxels per axis (a multiple of 8) */
    float origin[3];     /* cube centre, mm */
    float dim[3];        /* cube half-extent per axis, mm */
    int width, height;   /* depth / ray-cast image size */
    float min_dist, max_dist;  /* depth window, mm */
} rt_tsdf_params;

/* The reference's grid: 256^3 voxels, centre (0, 0.01, -1600.01), half-extent
 * 1500 mm (VoxelGrid::VoxelGrid, reconstruct/VoxelGrid.cpp:45-90,
 * rgbdscan/src/main.cpp:78), a 320x240 image, depth window 300..7000 mm. */
void rt_tsdf_default_params(rt_tsdf_params* p);
int rt_tsdf_create(int device, const rt_tsdf_params* p, rt_tsdf** out);
void rt_tsdf_destroy(rt_tsdf* g);
/* every voxel to (distance 1, weight 0) (resetKernel, VoxelGrid_cuda.cu:111-116) */
int rt_tsdf_reset(rt_tsdf* g);

/* The distance map the next update uses. From the reference's float4 RGBD
 * image: z = w (max - min) + min where 1e-4 < w < 1 - 1e-4, else 0, rows read
 * bottom-up (reconstructionKernel, VoxelGrid_cuda.cu:1088-1110); or from
 * 16-bit depth in mm (0 = no measurement), this library's input format. */
int rt_tsdf_set_depth_rgbd(rt_tsdf* g, const float* rgbd);
int rt_tsdf_set_depth_u16(rt_tsdf* g, const uint16_t* depth);

/* Fuse the distance map seen from camera pose T (updateVoxelsKernel,
 * VoxelGrid_cuda.cu:944-1014; VoxelGrid::update, VoxelGrid.cpp:284-298):
 * each voxel centre is projected into the map; where the four neighbouring
 * distances are all valid (>= one voxel) and within dim/200 of their bilinear
 * blend d, sdf = (d - voxel depth) / (0.03 dim); sdf < -1.25 is skipped,
 * otherwise clamp(sdf, -1, 1) enters the running mean with weight 1, the
 * weight capped at 16. */
int rt_tsdf_update(rt_tsdf* g, const float* T, const float* K);

/* Ray cast from camera pose T (rayCastVoxelsKernel + shadingKernel,
 * VoxelGrid_cuda.cu:118-164, 558-770, 844-894): per pixel the ray through the
 * normalised pixel position enters the cube, steps voxel by voxel to the
 * first negative distance, and the zero crossing is found by linear
 * interpolation of trilinear samples (three more refinements when `refine`).
 * xyz (float4 per pixel, camera coordinates, w = 1 inside the depth window)
 * and normals (float4, camera coordinates, from the neighbouring points, zero
 * where facing away or near the border); shade (float4, the normal's z as a
 * gray level) may be NULL. A voxel walk that would leave the grid ends the
 * ray (the reference reads outside the grid there). */
int rt_tsdf_raycast(rt_tsdf* g, const float* T, const float* K, int refine, float* xyz, float* normals, float* shade);

/* Resample src into dst under pose T (resampleVoxelsKernel,
 * VoxelGrid_cuda.cu:897-939; VoxelGrid::resample, VoxelGrid.cpp:362-369, with
 * T as setCameraParams hands it over): every dst voxel centre p is mapped to
 * T p and, where that lies within src's interpolation range [0, res-2] on
 * every axis, takes src's trilinear distance and weight there, else
 * (distance 1, weight 0). Both grids must have the same resolution, origin
 * and extent; dst != src. */
int rt_tsdf_resample(rt_tsdf* dst, const rt_tsdf* src, const float* T);

/* The grid as (distance, weight) float pairs, resolution^3 of them, x fastest
 * (host buffer). */
int rt_tsdf_read(rt_tsdf* g, float* host, size_t bytes);
/* Device pointer of the grid (the same layout), for callers that inspect it
 * on the device. */
const float* rt_tsdf_data(rt_tsdf* g);
/* The grid's stream (a hipStream_t) and a wait for everything queued on it. */
void* rt_tsdf_stream(rt_tsdf* g);
int rt_tsdf_synchronize(rt_tsdf* g);

#ifdef __cplusplus
}
#endif

#endif /* RT_TSDF_H */
