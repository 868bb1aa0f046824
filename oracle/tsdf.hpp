// CPU oracle for TSDF fusion and ray casting (the reference's VoxelGrid).
//
// TEST INFRASTRUCTURE ONLY: the checker of the HIP TSDF kernels; nothing in
// the product library links or calls it.
//
// A literal FP32 restatement of tracker/kernels/VoxelGrid_cuda.cu, evaluated
// voxel by voxel and ray by ray in the reference's operation order (compiled
// with -ffp-contract=off, as the device code is). Float -> unsigned
// conversions follow the device's saturating rule (negative and NaN -> 0).
// Where the reference would read outside the grid, the restatement bounds the
// read, and the HIP kernels apply the same bounds:
//   - trilinear samples: an address past the last voxel is held at the last
//     voxel (the reference reads beyond the array);
//   - the voxel walk ends the ray when the voxel index leaves the grid (the
//     reference reads outside it or, below zero, stays at index 0);
//   - the backward zero search takes at most 4*res+16 half-voxel steps (the
//     reference's loop ends when t0 passes the entry point, which these steps
//     reach unless t is so large that a half voxel is below its precision).
// These edges are "parity unpinned" against the reference; everything else is
// pinned by closed forms (planes and spheres fused and ray cast, see the tests).
#pragma once
#include <cfloat>
#include <cmath>
#include <cstddef>
#include <cstdint>

#include "oracle.hpp"

namespace oracle {
namespace tsdf {

struct Grid {
    int res, w, h;
    float origin[3], dim[3], vsize[3];  // centre, half-extent, half voxel edge
    float minDist, maxDist;
};

inline void make_grid(Grid& g, int res, const float* origin, const float* dim, int w, int h, float mn, float mx) {
    g.res = res; g.w = w; g.h = h;
    for (int i = 0; i < 3; i++) {
        g.origin[i] = origin[i];
        g.dim[i] = dim[i];
        g.vsize[i] = dim[i] / res;  // VoxelGrid.cpp:106
    }
    g.minDist = mn; g.maxDist = mx;
}

inline unsigned to_u32(float x) { return (x > 0.0f) ? (unsigned)x : 0u; }  // x < 2^32 in every use here

// reconstructionKernel (VoxelGrid_cuda.cu:1088-1110)
inline void dist_rgbd(const Grid& g, const float* rgbd, float* dist) {
    for (int y = 0; y < g.h; y++)
        for (int x = 0; x < g.w; x++) {
            const float wv = rgbd[4 * ((size_t)x + (size_t)(g.h - 1 - y) * g.w) + 3];
            const float tol = 1e-4f;
            dist[(size_t)x + (size_t)y * g.w] =
                (wv > tol && wv < (1.0f - tol)) ? wv * (g.maxDist - g.minDist) + g.minDist : 0.0f;
        }
}

// updateVoxelsKernel (VoxelGrid_cuda.cu:944-1014) with T = inverse of the
// camera pose (VoxelGrid::setCameraParams, VoxelGrid.cpp:280-286)
inline void update(const Grid& g, float* tsdf, const float* dist, const float* Tcur, const float* K) {
    float T[16];
    invert_rt4<float>(Tcur, T);
    const int w = g.w, h = g.h, res = g.res;
    float vs[3] = {g.vsize[0], g.vsize[1], g.vsize[2]};
    const float start[3] = {g.origin[0] - g.dim[0] + vs[0], g.origin[1] - g.dim[1] + vs[1], g.origin[2] - g.dim[2] + vs[2]};
    vs[0] *= 2; vs[1] *= 2; vs[2] *= 2;
    for (unsigned iz = 0; iz < (unsigned)res; iz++)
        for (unsigned iy = 0; iy < (unsigned)res; iy++)
            for (unsigned ix = 0; ix < (unsigned)res; ix++) {
                const size_t idx = ix + (size_t)iy * res + (size_t)iz * res * res;
                const float px = start[0] + ix * vs[0], py = start[1] + iy * vs[1], pz = start[2] + iz * vs[2];
                float rx = T[0] * px + T[1] * py + T[2] * pz + T[3];
                float ry = T[4] * px + T[5] * py + T[6] * pz + T[7];
                float rz = T[8] * px + T[9] * py + T[10] * pz + T[11];
                const float voxelDist = -rz;
                rx /= rz; ry /= rz;
                const float u = (rx * K[0] + K[2]) * (float)w;
                const float v = (ry * K[4] + K[5]) * (float)h;
                if (!(std::fabs(u) < 1e9f && std::fabs(v) < 1e9f)) continue;  // int() of these is rejected below
                const int ui = (int)u, vi = (int)v;
                if (ui < 1 || ui >= (w - 1 - 1) || vi < 1 || vi >= (h - 1 - 1)) continue;
                const size_t off = (size_t)ui + (size_t)vi * w;
                const float m3a = dist[off], m3b = dist[off + 1], m3c = dist[off + w], m3d = dist[off + w + 1];
                if (m3a < g.vsize[0] * 2 || m3b < g.vsize[0] * 2 || m3c < g.vsize[0] * 2 || m3d < g.vsize[0] * 2) continue;
                const float fu = u - ui, fv = v - vi;
                const float d = ((1 - fu) * (1 - fv) * m3a + fu * (1 - fv) * m3b + (1 - fu) * fv * m3c + fu * fv * m3d);
                const float lim = g.dim[0] / 200.0f;
                if (std::fabs(d - m3a) > lim || std::fabs(d - m3b) > lim || std::fabs(d - m3c) > lim ||
                    std::fabs(d - m3d) > lim)
                    continue;
                const float slopeRange = g.dim[0] * 0.030f;
                const float sdf = -(voxelDist - d) / slopeRange;
                if (sdf < -5.0 / 4.0f) continue;
                const float newTsdf = std::fmax(std::fmin(1.0f, sdf), -1.0f);
                float* t = tsdf + 2 * idx;
                t[0] = (t[1] * t[0] + 1.0f * newTsdf) / (t[1] + 1.0f);
                t[1] = std::fmin(1.0f + t[1], 16.0f);
            }
}

// trilinearDensity (VoxelGrid_cuda.cu:181-222)
inline float trilinear(const Grid& g, const float* tsdf, float px, float py, float pz) {
    const unsigned lim = (unsigned)g.res + 1;  // past the grid: the reads below are held at its last voxel
    auto index = [&](float p) { return p > 0.0f ? std::min((unsigned)std::fmin(p, (float)lim), lim) : 0u; };
    const unsigned ix = index(px), iy = index(py), iz = index(pz);
    const float fx = px - ix, fy = py - iy, fz = pz - iz;
    const size_t ys = g.res, zs = (size_t)g.res * g.res, last = zs * g.res - 1;
    const size_t o = ix + iy * ys + iz * zs;
    auto at = [&](size_t a) { return tsdf[2 * std::min(a, last)]; };
    const float d000 = at(o), d001 = at(o + zs), d010 = at(o + ys), d011 = at(o + ys + zs);
    const float d100 = at(o + 1), d101 = at(o + 1 + zs), d110 = at(o + 1 + ys), d111 = at(o + 1 + ys + zs);
    const float x00 = (1 - fx) * d000 + fx * d100;
    const float x01 = (1 - fx) * d001 + fx * d101;
    const float x10 = (1 - fx) * d010 + fx * d110;
    const float x11 = (1 - fx) * d011 + fx * d111;
    const float z0 = (1 - fy) * x00 + fy * x10;
    const float z1 = (1 - fy) * x01 + fy * x11;
    return (1 - fz) * z0 + fz * z1;
}

// resampleVoxelsKernel with trilinearDensity / trilinearWeight
// (VoxelGrid_cuda.cu:181-262, 897-939): iT maps dst voxel centres into src
inline void resample(const Grid& g, float* dst, const float* src, const float* iT) {
    const int res = g.res;
    float vs[3] = {g.vsize[0], g.vsize[1], g.vsize[2]};
    const float start[3] = {g.origin[0] - g.dim[0] + vs[0], g.origin[1] - g.dim[1] + vs[1], g.origin[2] - g.dim[2] + vs[2]};
    vs[0] *= 2; vs[1] *= 2; vs[2] *= 2;
    for (unsigned iz = 0; iz < (unsigned)res; iz++)
        for (unsigned iy = 0; iy < (unsigned)res; iy++)
            for (unsigned ix = 0; ix < (unsigned)res; ix++) {
                const size_t idx = ix + (size_t)iy * res + (size_t)iz * res * res;
                const float px = start[0] + ix * vs[0], py = start[1] + iy * vs[1], pz = start[2] + iz * vs[2];
                const float r[3] = {iT[0] * px + iT[1] * py + iT[2] * pz + iT[3], iT[4] * px + iT[5] * py + iT[6] * pz + iT[7],
                                    iT[8] * px + iT[9] * py + iT[10] * pz + iT[11]};
                float iv[3];
                bool out = false;
                for (int k = 0; k < 3; k++) {
                    iv[k] = (r[k] - start[k]) / vs[k];
                    if (!(iv[k] >= 0 && iv[k] <= res - 2)) out = true;  // NaN counts as outside
                }
                if (out) {
                    dst[2 * idx] = 1;
                    dst[2 * idx + 1] = 0;
                    continue;
                }
                for (int c = 0; c < 2; c++) {  // density, then weight
                    const unsigned x = (unsigned)iv[0], y = (unsigned)iv[1], z = (unsigned)iv[2];
                    const float fx = iv[0] - x, fy = iv[1] - y, fz = iv[2] - z;
                    const size_t ys = res, zs = (size_t)res * res, o = x + y * ys + z * zs;
                    auto at = [&](size_t a) { return src[2 * a + c]; };
                    const float x00 = (1 - fx) * at(o) + fx * at(o + 1);
                    const float x01 = (1 - fx) * at(o + zs) + fx * at(o + 1 + zs);
                    const float x10 = (1 - fx) * at(o + ys) + fx * at(o + 1 + ys);
                    const float x11 = (1 - fx) * at(o + ys + zs) + fx * at(o + 1 + ys + zs);
                    const float z0 = (1 - fy) * x00 + fy * x10;
                    const float z1 = (1 - fy) * x01 + fy * x11;
                    dst[2 * idx + c] = (1 - fz) * z0 + fz * z1;
                }
            }
}

inline float clampf(float x, float a, float b) { return std::fmin(std::fmax(x, a), b); }

// estimateLinearZero (VoxelGrid_cuda.cu:558-561)
inline float linear_zero(float tA, float dA, float tB, float dB) {
    const float s = clampf(dA / (dA - dB), 0.0f, 1.0f);
    return tA + s * (tB - tA);
}

struct V3 {
    float x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// traverseVoxelsCuda (VoxelGrid_cuda.cu:647-739): true and *tHit on a zero crossing
inline bool traverse(const Grid& g, const float* tsdf, V3 origin, V3 dir, V3 invDir, const int* sgn, float enterT,
                     float exitT, V3 ib, V3 vd, bool filtering, float* tHit) {
    enterT += vd.x * 2;
    exitT -= vd.x * 2;
    if (enterT < 0) enterT = 0;
    if (enterT + vd.x * 2 > exitT) return false;
    const V3 ivd = {1.0f / vd.x, 1.0f / vd.y, 1.0f / vd.z};
    const V3 q = (origin + enterT * dir - ib) * ivd;  // voxelize
    long vx = to_u32(q.x), vy = to_u32(q.y), vz = to_u32(q.z);
    V3 vp = ib + V3{(float)vx, (float)vy, (float)vz} * vd;
    const V3 sign = {(float)(1 - 2 * sgn[0]), (float)(1 - 2 * sgn[1]), (float)(1 - 2 * sgn[2])};
    const V3 step = {sign.x * vd.x, sign.y * vd.y, sign.z * vd.z};
    V3 tNext = (vp + step - origin) * invDir;
    const V3 tStep = step * invDir;
    int iter = filtering ? 3 : 0;
    const long res = g.res;
    float t = enterT;
    auto sample = [&](float tt) { const V3 p = (origin + tt * dir - ib) * ivd; return trilinear(g, tsdf, p.x, p.y, p.z); };
    while (t < exitT) {
        if (vx < 0 || vy < 0 || vz < 0 || vx >= res || vy >= res || vz >= res) return false;
        const size_t offset = (size_t)(vx + vy * res + vz * res * res);
        if (tsdf[2 * offset] < 0.0f) {
            float t1 = clampf(dot(vp - origin, dir), enterT, exitT);
            float e1 = sample(t1);
            if (e1 < 0.0f) {
                float t0 = t1 - vd.x / 2;
                for (int back = 0; t0 > enterT && back < 4 * g.res + 16; back++) {
                    float e0 = sample(t0);
                    if (e0 > 0.0f) {
                        float tx = linear_zero(t0, e0, t1, e1);
                        while (iter > 0) {
                            const float ex = sample(tx);
                            if (ex < 0.0f) { t1 = tx; e1 = ex; } else { t0 = tx; e0 = ex; }
                            tx = linear_zero(t0, e0, t1, e1);
                            iter--;
                        }
                        *tHit = tx;
                        return true;
                    } else if (e0 > e1) {
                        t1 = t0;
                        e1 = e0;
                    }
                    t0 -= vd.x * 0.5f;
                }
                return false;
            }
        }
        if (tNext.x < tNext.y) {
            if (tNext.x < tNext.z) { t = tNext.x; vp.x += step.x; vx += (long)sign.x; tNext.x += tStep.x; }
            else { t = tNext.z; vp.z += step.z; vz += (long)sign.z; tNext.z += tStep.z; }
        } else {
            if (tNext.y < tNext.z) { t = tNext.y; vp.y += step.y; vy += (long)sign.y; tNext.y += tStep.y; }
            else { t = tNext.z; vp.z += step.z; vz += (long)sign.z; tNext.z += tStep.z; }
        }
    }
    return false;
}

// rayCastVoxelsKernel with initRay and intersectVoxelGridCuda
// (VoxelGrid_cuda.cu:118-164, 741-770); Tcur is the camera pose
// (VoxelGrid::rayCastImage hands it over as iT, VoxelGrid.cpp:300-306)
inline void raycast(const Grid& g, const float* tsdf, const float* T, const float* K, bool filtering, float* xyz) {
    const V3 b0 = {g.origin[0] - g.dim[0], g.origin[1] - g.dim[1], g.origin[2] - g.dim[2]};
    const V3 b1 = {g.origin[0] + g.dim[0], g.origin[1] + g.dim[1], g.origin[2] + g.dim[2]};
    const V3 ib = {g.origin[0] - g.dim[0] + g.vsize[0], g.origin[1] - g.dim[1] + g.vsize[1],
                   g.origin[2] - g.dim[2] + g.vsize[2]};
    const V3 vd = {2.0f * g.vsize[0], 2.0f * g.vsize[1], 2.0f * g.vsize[2]};
    for (int yi = 0; yi < g.h; yi++)
        for (int xi = 0; xi < g.w; xi++) {
            float* px = xyz + 4 * ((size_t)xi + (size_t)yi * g.w);
            const float pdx = (float)xi / (float)g.w, pdy = (float)yi / (float)g.h;
            V3 cam = {-(pdx - K[2]) / K[0], -(pdy - K[5]) / K[4], -1};
            float len = 0.0f;
            len += cam.x * cam.x;
            len += cam.y * cam.y;
            len += cam.z * cam.z;
            len = 1.0f / std::sqrt(len);
            cam.x *= len; cam.y *= len; cam.z *= len;
            const V3 dir = {T[0] * cam.x + T[1] * cam.y + T[2] * cam.z, T[4] * cam.x + T[5] * cam.y + T[6] * cam.z,
                            T[8] * cam.x + T[9] * cam.y + T[10] * cam.z};
            const V3 org = {T[3], T[7], T[11]};
            const int sgn[3] = {dir.x < 0, dir.y < 0, dir.z < 0};
            const V3 inv = {std::fabs(dir.x) > 1e-6f ? 1.0f / dir.x : (1 - 2 * sgn[0]) * FLT_MAX,
                            std::fabs(dir.y) > 1e-6f ? 1.0f / dir.y : (1 - 2 * sgn[1]) * FLT_MAX,
                            std::fabs(dir.z) > 1e-6f ? 1.0f / dir.z : (1 - 2 * sgn[2]) * FLT_MAX};
            const V3 bnd[2] = {b0, b1};
            const float x0 = (bnd[sgn[0]].x - org.x) * inv.x, x1 = (bnd[1 - sgn[0]].x - org.x) * inv.x;
            const float y0 = (bnd[sgn[1]].y - org.y) * inv.y, y1 = (bnd[1 - sgn[1]].y - org.y) * inv.y;
            const float z0 = (bnd[sgn[2]].z - org.z) * inv.z, z1 = (bnd[1 - sgn[2]].z - org.z) * inv.z;
            const float tmin = std::fmax(x0, std::fmax(y0, z0)), tmax = std::fmin(x1, std::fmin(y1, z1));
            float tHit;
            if (!(tmin >= tmax || tmax <= 0) && traverse(g, tsdf, org, dir, inv, sgn, tmin, tmax, ib, vd, filtering, &tHit)) {
                const V3 ip = tHit * cam;
                px[0] = ip.x; px[1] = ip.y; px[2] = ip.z;
                px[3] = (tHit > g.minDist && tHit < g.maxDist) ? 1.0f : 0.0f;
            } else {
                px[0] = px[1] = px[2] = px[3] = 0.0f;
            }
        }
}

// shadingKernel (VoxelGrid_cuda.cu:844-894)
inline void shade(const Grid& g, const float* xyz, float* image, float* normals) {
    const int w = g.w, h = g.h;
    for (int yi = 0; yi < h; yi++)
        for (int xi = 0; xi < w; xi++) {
            const size_t off = (size_t)xi + (size_t)yi * w;
            float* pix = image + 4 * off;
            float* nrm = normals + 4 * off;
            pix[0] = pix[1] = pix[2] = 0.0f; pix[3] = 1.0f;
            nrm[0] = nrm[1] = nrm[2] = nrm[3] = 0.0f;
            if (xi >= w - 3 || yi >= h - 3 || xi < 3 || yi < 3) continue;
            const float* p1x = xyz + 4 * (off + 1);
            const float* p0x = xyz + 4 * (off - 1);
            const float* p1y = xyz + 4 * (off + w);
            const float* p0y = xyz + 4 * (off - w);
            if (p1x[3] <= 0.0f || p0x[3] <= 0.0f || p1y[3] <= 0.0f || p0y[3] <= 0.0f) continue;
            const V3 u = {p1x[0] - p0x[0], p1x[1] - p0x[1], p1x[2] - p0x[2]};
            const V3 v = {p0y[0] - p1y[0], p0y[1] - p1y[1], p0y[2] - p1y[2]};
            V3 n = {u.y * v.z - u.z * v.y, -(u.x * v.z - u.z * v.x), u.x * v.y - u.y * v.x};
            float len = 0.0f;
            len += n.x * n.x;
            len += n.y * n.y;
            len += n.z * n.z;
            len = 1.0f / std::sqrt(len);
            n.x *= len; n.y *= len; n.z *= len;
            const float color = dot(n, V3{0.0f, 0.0f, 1.0f});
            if (color > 0.0f) {
                pix[0] = pix[1] = pix[2] = color;
                nrm[0] = n.x; nrm[1] = n.y; nrm[2] = n.z;
            }
        }
}

}  // namespace tsdf
}  // namespace oracle
