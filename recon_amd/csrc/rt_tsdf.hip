// Truncated signed distance fusion into a voxel grid and ray casting of the
// grid (the reference's VoxelGrid, kernels/VoxelGrid_cuda.cu). Every FP32
// value follows the reference's operation order (the library is built with
// -ffp-contract=off).
#include <hip/hip_runtime.h>

#include <float.h>
#include <math.h>
#include <string.h>

#include <vector>

#include "rt_host.hpp"
#include "rt_tsdf.h"

// ray-cast workgroup: RAY_TX x RAY_TY pixels (one ray each). One wave per
// workgroup spreads a 320x240 image's 1200 waves evenly over the CUs (a sweep
// of 16x16, 8x8, 16x4, 8x4 and 32x8: 145, 139, 133, 147, 143 us per frame)
#ifndef RAY_TX
#define RAY_TX 16
#endif
#ifndef RAY_TY
#define RAY_TY 4
#endif
#ifndef TSDF_AHEAD
#define TSDF_AHEAD 8  // voxels of a ray whose loads are in flight together
#endif

namespace rt {

struct TsdfArgs {
    int res, w, h;
    float ox, oy, oz;     // cube centre
    float dx, dy, dz;     // half-extent
    float vx, vy, vz;     // dim / res: half the voxel edge
    float mn, mx;
    float T[16];          // update: world -> camera; ray cast: camera -> world
    float K[9];
};

// updateVoxelsKernel (VoxelGrid_cuda.cu:944-1014)
// One thread per voxel, a wave along a row of x so that the grid is read and
// written in 512-byte runs; the voxel is read only where it is updated.
__global__ __launch_bounds__(256) void k_tsdf_update(float2* __restrict__ tsdf, const float* __restrict__ dist, TsdfArgs a) {
    const unsigned ix = blockIdx.x * 64 + (threadIdx.x & 63), iy = blockIdx.y * 4 + (threadIdx.x >> 6), iz = blockIdx.z;
    if (ix > (unsigned)a.res - 1 || iy > (unsigned)a.res - 1 || iz > (unsigned)a.res - 1) return;
    const size_t idx = ix + (size_t)iy * a.res + (size_t)iz * a.res * a.res;
    const float* K = a.K;
    const float* T = a.T;
    const int w = a.w, h = a.h;
    const float sx = a.ox - a.dx + a.vx, sy = a.oy - a.dy + a.vy, sz = a.oz - a.dz + a.vz;  // first voxel centre
    const float ex = a.vx * 2, ey = a.vy * 2, ez = a.vz * 2;                                // voxel edge
    const float px = sx + ix * ex, py = sy + iy * ey, pz = sz + iz * ez;
    float rx = T[0] * px + T[1] * py + T[2] * pz + T[3];
    float ry = T[4] * px + T[5] * py + T[6] * pz + T[7];
    const float rz = T[8] * px + T[9] * py + T[10] * pz + T[11];
    const float voxelDist = -rz;
    rx /= rz; ry /= rz;
    const float u = (rx * K[0] + K[2]) * (float)w, v = (ry * K[4] + K[5]) * (float)h;
    // a projection far outside the image (or not a number) is rejected below
    // in the reference as well; its int conversion is not defined
    if (!(fabsf(u) < 1e9f && fabsf(v) < 1e9f)) return;
    const int ui = (int)u, vi = (int)v;
    if (ui < 1 || ui >= w - 2 || vi < 1 || vi >= h - 2) return;
    const size_t off = (size_t)ui + (size_t)vi * w;
    const float m3a = dist[off], m3b = dist[off + 1], m3c = dist[off + w], m3d = dist[off + w + 1];
    const float vmin = a.vx * 2;
    if (m3a < vmin || m3b < vmin || m3c < vmin || m3d < vmin) return;
    const float fu = u - ui, fv = v - vi;
    const float d = ((1 - fu) * (1 - fv) * m3a + fu * (1 - fv) * m3b + (1 - fu) * fv * m3c + fu * fv * m3d);
    const float tol = a.dx / 200.0f;
    if (fabsf(d - m3a) > tol || fabsf(d - m3b) > tol || fabsf(d - m3c) > tol || fabsf(d - m3d) > tol) return;
    const float slope = a.dx * 0.030f;
    const float sdf = -(voxelDist - d) / slope;
    if (sdf < -1.25f) return;
    const float nt = fmaxf(fminf(1.0f, sdf), -1.0f);
    float2 t = tsdf[idx];
    t.x = (t.y * t.x + 1.0f * nt) / (t.y + 1.0f);
    t.y = fminf(1.0f + t.y, 16.0f);
    tsdf[idx] = t;
}

__global__ __launch_bounds__(256) void k_tsdf_reset(float2* __restrict__ tsdf, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) tsdf[i] = make_float2(1.0f, 0.0f);
}

// reconstructionKernel (VoxelGrid_cuda.cu:1088-1110): rows read bottom-up
__global__ __launch_bounds__(256) void k_tsdf_dist_rgbd(const float4* __restrict__ rgbd, int w, int h, float mn, float mx,
                                                        float* __restrict__ dist) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= w || y >= h) return;
    const float wv = rgbd[(size_t)x + (size_t)(h - 1 - y) * w].w;
    const float tol = 1e-4f;
    dist[(size_t)x + (size_t)y * w] = (wv > tol && wv < (1.0f - tol)) ? wv * (mx - mn) + mn : 0.0f;
}
__global__ __launch_bounds__(256) void k_tsdf_dist_u16(const uint16_t* __restrict__ d, size_t n, float* __restrict__ dist) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dist[i] = (float)d[i];
}

// trilinearDensity (VoxelGrid_cuda.cu:181-222) at grid coordinates p. The
// cell index and the eight addresses are the reference's; an address past the
// grid's last voxel (the reference reads beyond the array there) is held at
// the last voxel.
__device__ __forceinline__ float trilinear(const float2* __restrict__ g, int res, float px, float py, float pz) {
    const unsigned lim = (unsigned)res + 1;
    const unsigned ix = px > 0.0f ? min((unsigned)fminf(px, (float)lim), lim) : 0u;
    const unsigned iy = py > 0.0f ? min((unsigned)fminf(py, (float)lim), lim) : 0u;
    const unsigned iz = pz > 0.0f ? min((unsigned)fminf(pz, (float)lim), lim) : 0u;
    const float fx = px - ix, fy = py - iy, fz = pz - iz;
    const size_t ys = res, zs = (size_t)res * res, last = zs * res - 1;
    const size_t o = ix + iy * ys + iz * zs;
    auto at = [&](size_t a) { return g[a < last ? a : last].x; };
    const float d000 = at(o), d001 = at(o + zs), d010 = at(o + ys), d011 = at(o + ys + zs);
    const float d100 = at(o + 1), d101 = at(o + 1 + zs), d110 = at(o + 1 + ys), d111 = at(o + 1 + ys + zs);
    const float x00 = (1 - fx) * d000 + fx * d100, x01 = (1 - fx) * d001 + fx * d101;
    const float x10 = (1 - fx) * d010 + fx * d110, x11 = (1 - fx) * d011 + fx * d111;
    const float z0 = (1 - fy) * x00 + fy * x10, z1 = (1 - fy) * x01 + fy * x11;
    return (1 - fz) * z0 + fz * z1;
}
// resampleVoxelsKernel with trilinearDensity / trilinearWeight
// (VoxelGrid_cuda.cu:181-262, 897-939): T maps dst voxel centres into src
__global__ __launch_bounds__(256) void k_tsdf_resample(float2* __restrict__ dst, const float2* __restrict__ src, TsdfArgs a) {
    const unsigned ix = blockIdx.x * 64 + (threadIdx.x & 63), iy = blockIdx.y * 4 + (threadIdx.x >> 6), iz = blockIdx.z;
    if (ix > (unsigned)a.res - 1 || iy > (unsigned)a.res - 1 || iz > (unsigned)a.res - 1) return;
    const int res = a.res;
    const size_t idx = ix + (size_t)iy * res + (size_t)iz * res * res;
    const float* T = a.T;
    const float sx = a.ox - a.dx + a.vx, sy = a.oy - a.dy + a.vy, sz = a.oz - a.dz + a.vz;
    const float ex = a.vx * 2, ey = a.vy * 2, ez = a.vz * 2;
    const float px = sx + ix * ex, py = sy + iy * ey, pz = sz + iz * ez;
    const float rx = T[0] * px + T[1] * py + T[2] * pz + T[3];
    const float ry = T[4] * px + T[5] * py + T[6] * pz + T[7];
    const float rz = T[8] * px + T[9] * py + T[10] * pz + T[11];
    const float qx = (rx - sx) / ex, qy = (ry - sy) / ey, qz = (rz - sz) / ez;
    const float hi = (float)(res - 2);
    float2 out = make_float2(1.0f, 0.0f);
    // in range, so the cell index is at most res-2 on every axis and all eight reads are inside the grid
    if (qx >= 0 && qx <= hi && qy >= 0 && qy <= hi && qz >= 0 && qz <= hi) {
        const unsigned cx = (unsigned)qx, cy = (unsigned)qy, cz = (unsigned)qz;
        const float fx = qx - cx, fy = qy - cy, fz = qz - cz;
        const size_t ys = res, zs = (size_t)res * res;
        const size_t o = cx + cy * ys + cz * zs;
        const float2 d000 = src[o], d001 = src[o + zs], d010 = src[o + ys], d011 = src[o + ys + zs];
        const float2 d100 = src[o + 1], d101 = src[o + 1 + zs], d110 = src[o + 1 + ys], d111 = src[o + 1 + ys + zs];
        float r2[2];
#pragma unroll
        for (int c = 0; c < 2; c++) {
            const float a000 = c ? d000.y : d000.x, a001 = c ? d001.y : d001.x, a010 = c ? d010.y : d010.x,
                        a011 = c ? d011.y : d011.x, a100 = c ? d100.y : d100.x, a101 = c ? d101.y : d101.x,
                        a110 = c ? d110.y : d110.x, a111 = c ? d111.y : d111.x;
            const float x00 = (1 - fx) * a000 + fx * a100, x01 = (1 - fx) * a001 + fx * a101;
            const float x10 = (1 - fx) * a010 + fx * a110, x11 = (1 - fx) * a011 + fx * a111;
            const float z0 = (1 - fy) * x00 + fy * x10, z1 = (1 - fy) * x01 + fy * x11;
            r2[c] = (1 - fz) * z0 + fz * z1;
        }
        out = make_float2(r2[0], r2[1]);
    }
    dst[idx] = out;
}

// estimateLinearZero (VoxelGrid_cuda.cu:558-561)
__device__ __forceinline__ float linear_zero(float tA, float dA, float tB, float dB) {
    const float s = fminf(fmaxf(dA / (dA - dB), 0.0f), 1.0f);
    return tA + s * (tB - tA);
}

// rayCastVoxelsKernel with initRay, intersectVoxelGridCuda and
// traverseVoxelsCuda (VoxelGrid_cuda.cu:118-164, 633-770)
__global__ __launch_bounds__(RAY_TX * RAY_TY) void k_tsdf_raycast(const float2* __restrict__ g, TsdfArgs a, int refine,
                                                                  float4* __restrict__ out) {
    const int xi = blockIdx.x * RAY_TX + (threadIdx.x % RAY_TX), yi = blockIdx.y * RAY_TY + (threadIdx.x / RAY_TX);
    if (xi >= a.w || yi >= a.h) return;
    const float* K = a.K;
    const float* T = a.T;
    const int res = a.res;
    // initRay: the normalised pixel position through the inverse intrinsics, unit length
    const float pdx = (float)xi / (float)a.w, pdy = (float)yi / (float)a.h;
    float cx = -(pdx - K[2]) / K[0], cy = -(pdy - K[5]) / K[4], cz = -1.0f;
    float len = 0.0f;
    len += cx * cx; len += cy * cy; len += cz * cz;
    len = 1.0f / sqrtf(len);
    cx *= len; cy *= len; cz *= len;
    const float dx = T[0] * cx + T[1] * cy + T[2] * cz, dy = T[4] * cx + T[5] * cy + T[6] * cz,
                dz = T[8] * cx + T[9] * cy + T[10] * cz;
    const float ox = T[3], oy = T[7], oz = T[11];
    const int s0 = dx < 0, s1 = dy < 0, s2 = dz < 0;
    const float ivx = fabsf(dx) > 1e-6f ? 1.0f / dx : (1 - 2 * s0) * FLT_MAX;
    const float ivy = fabsf(dy) > 1e-6f ? 1.0f / dy : (1 - 2 * s1) * FLT_MAX;
    const float ivz = fabsf(dz) > 1e-6f ? 1.0f / dz : (1 - 2 * s2) * FLT_MAX;
    float4 result = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    // intersectVoxelGridCuda: slab test against the cube
    const float bx[2] = {a.ox - a.dx, a.ox + a.dx}, by[2] = {a.oy - a.dy, a.oy + a.dy}, bz[2] = {a.oz - a.dz, a.oz + a.dz};
    const float x0 = (bx[s0] - ox) * ivx, x1 = (bx[1 - s0] - ox) * ivx;
    const float y0 = (by[s1] - oy) * ivy, y1 = (by[1 - s1] - oy) * ivy;
    const float z0 = (bz[s2] - oz) * ivz, z1 = (bz[1 - s2] - oz) * ivz;
    const float tmin = fmaxf(x0, fmaxf(y0, z0)), tmax = fminf(x1, fminf(y1, z1));
    if (!(tmin >= tmax || tmax <= 0)) {
        float enterT = tmin, exitT = tmax;
        const float vdx = 2.0f * a.vx, vdy = 2.0f * a.vy, vdz = 2.0f * a.vz;  // voxel edge
        const float ibx = a.ox - a.dx + a.vx, iby = a.oy - a.dy + a.vy, ibz = a.oz - a.dz + a.vz;
        enterT += vdx * 2; exitT -= vdx * 2;
        if (enterT < 0) enterT = 0;
        bool hit = false;
        float tHit = 0.0f;
        if (!(enterT + vdx * 2 > exitT)) {
            const float ivdx = 1.0f / vdx, ivdy = 1.0f / vdy, ivdz = 1.0f / vdz;
            // voxelize: the voxel the ray enters
            const float qx = (ox + enterT * dx - ibx) * ivdx, qy = (oy + enterT * dy - iby) * ivdy,
                        qz = (oz + enterT * dz - ibz) * ivdz;
            int vx = qx > 0.0f ? (int)qx : 0, vy = qy > 0.0f ? (int)qy : 0, vz = qz > 0.0f ? (int)qz : 0;
            float vpx = ibx + (float)vx * vdx, vpy = iby + (float)vy * vdy, vpz = ibz + (float)vz * vdz;
            const int gx = 1 - 2 * s0, gy = 1 - 2 * s1, gz = 1 - 2 * s2;
            const float stx = (float)gx * vdx, sty = (float)gy * vdy, stz = (float)gz * vdz;
            float tnx = (vpx + stx - ox) * ivx, tny = (vpy + sty - oy) * ivy, tnz = (vpz + stz - oz) * ivz;
            const float tsx = stx * ivx, tsy = sty * ivy, tsz = stz * ivz;
            int iter = refine ? 3 : 0;
            // The walk does not depend on the distances it reads, so TSDF_AHEAD
            // voxels are stepped and their loads issued together, then tested
            // in walk order: the same voxels, tests and arithmetic as one
            // voxel at a time, with the load latency paid once per group.
            const float* gd = reinterpret_cast<const float*>(g);
            float t = enterT;
            bool done = false;
            // the grid check ends the walk within 3 * (res + 1) steps: each step moves one index one way
            while (!done) {
                float val[TSDF_AHEAD], wpx[TSDF_AHEAD], wpy[TSDF_AHEAD], wpz[TSDF_AHEAD];
                bool live = true;
                int nlive = 0;
                // branch-free, so that the loads are issued back to back: each
                // reads a voxel clamped into the grid, and a voxel past the end
                // of the walk is ignored
#pragma unroll
                for (int j = 0; j < TSDF_AHEAD; j++) {
                    live = live && t < exitT && vx >= 0 && vy >= 0 && vz >= 0 && vx < res && vy < res && vz < res;
                    const size_t offset = (size_t)min(max(vx, 0), res - 1) + (size_t)min(max(vy, 0), res - 1) * res +
                                          (size_t)min(max(vz, 0), res - 1) * res * res;
                    val[j] = gd[2 * offset];
                    nlive += live ? 1 : 0;
                    wpx[j] = vpx; wpy[j] = vpy; wpz[j] = vpz;
                    // the step to the nearest voxel wall: x if tnx < tny and tnx < tnz, else y if
                    // tny <= tnx and tny < tnz, else z
                    const bool sx = tnx < tny && tnx < tnz, sy = !(tnx < tny) && tny < tnz, sz = !sx && !sy;
                    t = sx ? tnx : (sy ? tny : tnz);
                    vpx = sx ? vpx + stx : vpx; vx = sx ? vx + gx : vx; tnx = sx ? tnx + tsx : tnx;
                    vpy = sy ? vpy + sty : vpy; vy = sy ? vy + gy : vy; tny = sy ? tny + tsy : tny;
                    vpz = sz ? vpz + stz : vpz; vz = sz ? vz + gz : vz; tnz = sz ? tnz + tsz : tnz;
                }
                unsigned cand = 0;
#pragma unroll
                for (int j = 0; j < TSDF_AHEAD; j++) cand |= (j < nlive && val[j] < 0.0f) ? (1u << j) : 0u;
                while (cand && !done) {
                    const int j = __builtin_ctz(cand);
                    cand &= cand - 1;
                    float px = wpx[0], py = wpy[0], pz = wpz[0];
#pragma unroll
                    for (int k = 1; k < TSDF_AHEAD; k++)
                        if (k == j) { px = wpx[k]; py = wpy[k]; pz = wpz[k]; }
                    float t1 = fminf(fmaxf((px - ox) * dx + (py - oy) * dy + (pz - oz) * dz, enterT), exitT);
                    float e1 = trilinear(g, res, (ox + t1 * dx - ibx) * ivdx, (oy + t1 * dy - iby) * ivdy,
                                         (oz + t1 * dz - ibz) * ivdz);
                    if (e1 < 0.0f) {
                        done = true;  // found, or the surface was expected and not found
                        float t0 = t1 - vdx / 2;
                        // at most the cube's diagonal in half-voxel steps
                        for (int back = 0; t0 > enterT && back < 4 * res + 16; back++) {
                            float e0 = trilinear(g, res, (ox + t0 * dx - ibx) * ivdx, (oy + t0 * dy - iby) * ivdy,
                                                 (oz + t0 * dz - ibz) * ivdz);
                            if (e0 > 0.0f) {
                                float tx = linear_zero(t0, e0, t1, e1);
                                while (iter > 0) {
                                    const float exv = trilinear(g, res, (ox + tx * dx - ibx) * ivdx,
                                                                (oy + tx * dy - iby) * ivdy, (oz + tx * dz - ibz) * ivdz);
                                    if (exv < 0.0f) { t1 = tx; e1 = exv; } else { t0 = tx; e0 = exv; }
                                    tx = linear_zero(t0, e0, t1, e1);
                                    iter--;
                                }
                                tHit = tx;
                                hit = true;
                                break;
                            } else if (e0 > e1) {
                                t1 = t0; e1 = e0;
                            }
                            t0 -= vdx * 0.5f;
                        }
                    }
                }
                if (nlive < TSDF_AHEAD) done = true;  // the walk passed the exit or left the grid
            }
        }
        if (hit) result = make_float4(tHit * cx, tHit * cy, tHit * cz, (tHit > a.mn && tHit < a.mx) ? 1.0f : 0.0f);
    }
    out[(size_t)xi + (size_t)yi * a.w] = result;
}

// shadingKernel (VoxelGrid_cuda.cu:844-894): normals from the neighbouring points
__global__ __launch_bounds__(256) void k_tsdf_shade(const float4* __restrict__ xyz, int w, int h, float4* __restrict__ shade,
                                                    float4* __restrict__ normals) {
    const int xi = blockIdx.x * 16 + (threadIdx.x & 15), yi = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (xi >= w || yi >= h) return;
    const size_t off = (size_t)xi + (size_t)yi * w;
    float4 px = make_float4(0.0f, 0.0f, 0.0f, 1.0f), nm = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (!(xi >= w - 3 || yi >= h - 3 || xi < 3 || yi < 3)) {
        const float4 p1x = xyz[off + 1], p0x = xyz[off - 1], p1y = xyz[off + w], p0y = xyz[off - w];
        if (p1x.w > 0.0f && p0x.w > 0.0f && p1y.w > 0.0f && p0y.w > 0.0f) {
            const float ux = p1x.x - p0x.x, uy = p1x.y - p0x.y, uz = p1x.z - p0x.z;
            const float vx = p0y.x - p1y.x, vy = p0y.y - p1y.y, vz = p0y.z - p1y.z;
            float nx = uy * vz - uz * vy, ny = -(ux * vz - uz * vx), nz = ux * vy - uy * vx;
            float len = 0.0f;
            len += nx * nx; len += ny * ny; len += nz * nz;
            len = 1.0f / sqrtf(len);
            nx *= len; ny *= len; nz *= len;
            const float c = nx * 0.0f + ny * 0.0f + nz * 1.0f;
            if (c > 0.0f) {
                px = make_float4(c, c, c, 1.0f);
                nm = make_float4(nx, ny, nz, 0.0f);
            }
        }
    }
    if (shade) shade[off] = px;
    normals[off] = nm;
}

}  // namespace rt

using namespace rt;

struct rt_tsdf {
    int device = 0;
    rt_tsdf_params p;
    hipStream_t stream = nullptr;
    float2* grid = nullptr;
    float* dist = nullptr;
    hipEvent_t done = nullptr;  // orders work between two grids' streams
};

static TsdfArgs tsdf_args(const rt_tsdf* g, const float* T, const float* K) {
    TsdfArgs a;
    a.res = g->p.resolution; a.w = g->p.width; a.h = g->p.height;
    a.ox = g->p.origin[0]; a.oy = g->p.origin[1]; a.oz = g->p.origin[2];
    a.dx = g->p.dim[0]; a.dy = g->p.dim[1]; a.dz = g->p.dim[2];
    a.vx = a.dx / a.res; a.vy = a.dy / a.res; a.vz = a.dz / a.res;  // VoxelGrid.cpp:106
    a.mn = g->p.min_dist; a.mx = g->p.max_dist;
    memcpy(a.T, T, sizeof(a.T));
    memcpy(a.K, K, sizeof(a.K));
    return a;
}

extern "C" void rt_tsdf_default_params(rt_tsdf_params* p) {
    if (!p) return;
    p->resolution = 256;
    p->origin[0] = 0.0f; p->origin[1] = 0.01f; p->origin[2] = -1600.01f;
    p->dim[0] = p->dim[1] = p->dim[2] = 1500.0f;
    p->width = 320; p->height = 240;
    p->min_dist = 300.0f; p->max_dist = 7000.0f;
}

extern "C" int rt_tsdf_create(int device, const rt_tsdf_params* p, rt_tsdf** out) {
    if (!p || !out || p->resolution < 8 || p->resolution % 8 || p->resolution > 1024 || p->width < 8 || p->height < 8 ||
        !(p->dim[0] > 0 && p->dim[1] > 0 && p->dim[2] > 0))
        return set_error(RT_ERR_ARG, "rt_tsdf_create: bad argument");
    if (hipSetDevice(device) != hipSuccess) return set_error(RT_ERR_HIP, "hipSetDevice");
    rt_tsdf* g = new rt_tsdf();
    g->device = device; g->p = *p;
    const size_t nv = (size_t)p->resolution * p->resolution * p->resolution;
    if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess) { delete g; return set_error(RT_ERR_HIP, "stream"); }
    if (hipEventCreateWithFlags(&g->done, hipEventDisableTiming) != hipSuccess) {
        hipStreamDestroy(g->stream);
        delete g;
        return set_error(RT_ERR_HIP, "event");
    }
    if (hipMalloc((void**)&g->grid, nv * sizeof(float2)) != hipSuccess ||
        hipMalloc((void**)&g->dist, (size_t)p->width * p->height * sizeof(float)) != hipSuccess) {
        if (g->grid) hipFree(g->grid);
        hipEventDestroy(g->done);
        hipStreamDestroy(g->stream);
        delete g;
        return set_error(RT_ERR_NOMEM, "rt_tsdf_create: grid");
    }
    if (hipMemsetAsync(g->dist, 0, (size_t)p->width * p->height * sizeof(float), g->stream) != hipSuccess) {
        rt_tsdf_destroy(g);
        return set_error(RT_ERR_HIP, "rt_tsdf_create: clear");
    }
    int rc = rt_tsdf_reset(g);
    if (rc) { rt_tsdf_destroy(g); return rc; }
    *out = g;
    return RT_OK;
}

extern "C" void rt_tsdf_destroy(rt_tsdf* g) {
    if (!g) return;
    hipSetDevice(g->device);
    hipStreamSynchronize(g->stream);
    hipFree(g->grid);
    hipFree(g->dist);
    hipEventDestroy(g->done);
    hipStreamDestroy(g->stream);
    delete g;
}

extern "C" int rt_tsdf_reset(rt_tsdf* g) {
    if (!g) return set_error(RT_ERR_ARG, "rt_tsdf_reset: null");
    if (hipSetDevice(g->device) != hipSuccess) return set_error(RT_ERR_HIP, "hipSetDevice");
    const size_t n = (size_t)g->p.resolution * g->p.resolution * g->p.resolution;
    k_tsdf_reset<<<(unsigned)((n + 255) / 256), 256, 0, g->stream>>>(g->grid, n);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(g->stream) != hipSuccess)
        return set_error(RT_ERR_HIP, "rt_tsdf_reset");
    return RT_OK;
}

extern "C" int rt_tsdf_set_depth_rgbd(rt_tsdf* g, const float* rgbd) {
    if (!g || !rgbd) return set_error(RT_ERR_ARG, "rt_tsdf_set_depth_rgbd: null");
    if (hipSetDevice(g->device) != hipSuccess) return set_error(RT_ERR_HIP, "hipSetDevice");
    dim3 grid((g->p.width + 15) / 16, (g->p.height + 15) / 16);
    k_tsdf_dist_rgbd<<<grid, 256, 0, g->stream>>>((const float4*)rgbd, g->p.width, g->p.height, g->p.min_dist,
                                                   g->p.max_dist, g->dist);
    if (hipGetLastError() != hipSuccess) return set_error(RT_ERR_HIP, "rt_tsdf_set_depth_rgbd");
    return RT_OK;
}

extern "C" int rt_tsdf_set_depth_u16(rt_tsdf* g, const uint16_t* depth) {
    if (!g || !depth) return set_error(RT_ERR_ARG, "rt_tsdf_set_depth_u16: null");
    if (hipSetDevice(g->device) != hipSuccess) return set_error(RT_ERR_HIP, "hipSetDevice");
    const size_t n = (size_t)g->p.width * g->p.height;
    k_tsdf_dist_u16<<<(unsigned)((n + 255) / 256), 256, 0, g->stream>>>(depth, n, g->dist);
    if (hipGetLastError() != hipSuccess) return set_error(RT_ERR_HIP, "rt_tsdf_set_depth_u16");
    return RT_OK;
}

extern "C" int rt_tsdf_update(rt_tsdf* g, const float* T, const float* K) {
    if (!g || !T || !K) return set_error(RT_ERR_ARG, "rt_tsdf_update: null");
    if (hipSetDevice(g->device) != hipSuccess) return set_error(RT_ERR_HIP, "hipSetDevice");
    float iT[16];
    invert_rt4f(T, iT);  // VoxelGrid::setCameraParams (VoxelGrid.cpp:275-282): the kernel maps world -> camera
    const TsdfArgs a = tsdf_args(g, iT, K);
    const unsigned r = (unsigned)g->p.resolution;
    k_tsdf_update<<<dim3((r + 63) / 64, r / 4, r), 256, 0, g->stream>>>(g->grid, g->dist, a);
    if (hipGetLastError() != hipSuccess) return set_error(RT_ERR_HIP, "rt_tsdf_update");
    return RT_OK;
}

extern "C" int rt_tsdf_raycast(rt_tsdf* g, const float* T, const float* K, int refine, float* xyz, float* normals,
                               float* shade) {
    if (!g || !T || !K || !xyz || !normals) return set_error(RT_ERR_ARG, "rt_tsdf_raycast: null");
    if (hipSetDevice(g->device) != hipSuccess) return set_error(RT_ERR_HIP, "hipSetDevice");
    const TsdfArgs a = tsdf_args(g, T, K);  // the ray cast walks from the camera: camera -> world
    dim3 grid((g->p.width + 15) / 16, (g->p.height + 15) / 16);
    dim3 rgrid((g->p.width + RAY_TX - 1) / RAY_TX, (g->p.height + RAY_TY - 1) / RAY_TY);
    k_tsdf_raycast<<<rgrid, RAY_TX * RAY_TY, 0, g->stream>>>(g->grid, a, refine ? 1 : 0, (float4*)xyz);
    k_tsdf_shade<<<grid, 256, 0, g->stream>>>((const float4*)xyz, g->p.width, g->p.height, (float4*)shade,
                                               (float4*)normals);
    if (hipGetLastError() != hipSuccess) return set_error(RT_ERR_HIP, "rt_tsdf_raycast");
    return RT_OK;
}

extern "C" int rt_tsdf_resample(rt_tsdf* dst, const rt_tsdf* src, const float* T) {
    if (!dst || !src || !T || dst == src) return set_error(RT_ERR_ARG, "rt_tsdf_resample: null or same grid");
    const rt_tsdf_params &a = dst->p, &b = src->p;
    if (a.resolution != b.resolution || memcmp(a.origin, b.origin, sizeof(a.origin)) || memcmp(a.dim, b.dim, sizeof(a.dim)) ||
        dst->device != src->device)
        return set_error(RT_ERR_ARG, "rt_tsdf_resample: grids differ in resolution, origin, extent or device");
    if (hipSetDevice(dst->device) != hipSuccess) return set_error(RT_ERR_HIP, "hipSetDevice");
    float K[9] = {0};
    const TsdfArgs args = tsdf_args(dst, T, K);  // VoxelGrid::setCameraParams keeps T itself as iT
    // src's queued work is ordered before the resample
    if (hipEventRecord(src->done, src->stream) != hipSuccess || hipStreamWaitEvent(dst->stream, src->done, 0) != hipSuccess)
        return set_error(RT_ERR_HIP, "rt_tsdf_resample");
    const unsigned r = (unsigned)a.resolution;
    k_tsdf_resample<<<dim3((r + 63) / 64, r / 4, r), 256, 0, dst->stream>>>(dst->grid, src->grid, args);
    if (hipGetLastError() != hipSuccess) return set_error(RT_ERR_HIP, "rt_tsdf_resample");
    // src's later work is ordered after the resample has read it
    if (hipEventRecord(dst->done, dst->stream) != hipSuccess || hipStreamWaitEvent(src->stream, dst->done, 0) != hipSuccess)
        return set_error(RT_ERR_HIP, "rt_tsdf_resample");
    return RT_OK;
}

extern "C" int rt_tsdf_read(rt_tsdf* g, float* host, size_t bytes) {
    if (!g || !host) return set_error(RT_ERR_ARG, "rt_tsdf_read: null");
    const size_t n = (size_t)g->p.resolution * g->p.resolution * g->p.resolution * sizeof(float2);
    if (bytes < n) return set_error(RT_ERR_ARG, "rt_tsdf_read: buffer too small");
    if (hipSetDevice(g->device) != hipSuccess || hipMemcpyAsync(host, g->grid, n, hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
        hipStreamSynchronize(g->stream) != hipSuccess)
        return set_error(RT_ERR_HIP, "rt_tsdf_read");
    return RT_OK;
}

extern "C" const float* rt_tsdf_data(rt_tsdf* g) { return g ? (const float*)g->grid : nullptr; }

extern "C" void* rt_tsdf_stream(rt_tsdf* g) { return g ? (void*)g->stream : nullptr; }

extern "C" int rt_tsdf_synchronize(rt_tsdf* g) {
    if (!g) return set_error(RT_ERR_ARG, "rt_tsdf_synchronize: null");
    if (hipSetDevice(g->device) != hipSuccess || hipStreamSynchronize(g->stream) != hipSuccess)
        return set_error(RT_ERR_HIP, "rt_tsdf_synchronize");
    return RT_OK;
}
