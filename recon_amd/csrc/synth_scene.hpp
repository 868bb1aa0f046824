// Seeded synthetic RGB-D scenes, shared by the CPU and the GPU renderer.
//
// A scene is a room of five slightly tilted planes (back wall, floor, ceiling,
// two side walls) with 2-5 randomly oriented boxes inside it. Surfaces carry a
// view-independent fractal value-noise texture (5 octaves) evaluated at the 3D
// hit point, so brightness constancy holds between views. A frame is ray cast
// from a camera pose: 8-bit gray from 2x2 supersamples, 16-bit depth in mm
// along the optical axis, valid in [500, 4000] mm, with i.i.d. Gaussian noise
// of a given fraction of depth (0 = no measurement). Camera frames: x right,
// y up, looking down -z; poses are row-major 4x4 doubles T_wc.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace synth {

constexpr int MAX_BOXES = 5;

struct Plane { double n[3], c, albedo, wavelength; uint32_t tseed; };
struct Box { double ctr[3], R[9], h[3], albedo, wavelength; uint32_t tseed; };
struct Scene {
    Plane planes[5];
    Box boxes[MAX_BOXES];
    int nboxes;
};

struct Rng {  // splitmix64 stream
    uint64_t s;
    __host__ __device__ explicit Rng(uint64_t seed) : s(seed ^ 0x9E3779B97F4A7C15ull) {}
    __host__ __device__ uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    __host__ __device__ double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    __host__ __device__ double uni(double a, double b) { return a + (b - a) * uni(); }
};

__host__ __device__ inline uint32_t hash3(int x, int y, int z, uint32_t seed) {
    uint32_t h = seed * 0x9E3779B1u;
    h ^= (uint32_t)x * 0x85EBCA77u; h = (h << 13) | (h >> 19); h *= 0xC2B2AE3Du;
    h ^= (uint32_t)y * 0x27D4EB2Fu; h = (h << 13) | (h >> 19); h *= 0x165667B1u;
    h ^= (uint32_t)z * 0x9E3779B1u; h = (h << 13) | (h >> 19); h *= 0x85EBCA77u;
    h ^= h >> 16; h *= 0x7FEB352Du; h ^= h >> 15; h *= 0x846CA68Bu; h ^= h >> 16;
    return h;
}
__host__ __device__ inline double lattice(int x, int y, int z, uint32_t seed) {
    return (double)(hash3(x, y, z, seed) >> 8) * (2.0 / 16777216.0) - 1.0;
}
__host__ __device__ inline double fade(double t) { return t * t * (3.0 - 2.0 * t); }
__host__ __device__ inline double value_noise(double x, double y, double z, uint32_t seed) {
    double fx = floor(x), fy = floor(y), fz = floor(z);
    int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    double tx = fade(x - fx), ty = fade(y - fy), tz = fade(z - fz);
    double c000 = lattice(ix, iy, iz, seed), c100 = lattice(ix + 1, iy, iz, seed);
    double c010 = lattice(ix, iy + 1, iz, seed), c110 = lattice(ix + 1, iy + 1, iz, seed);
    double c001 = lattice(ix, iy, iz + 1, seed), c101 = lattice(ix + 1, iy, iz + 1, seed);
    double c011 = lattice(ix, iy + 1, iz + 1, seed), c111 = lattice(ix + 1, iy + 1, iz + 1, seed);
    double x00 = c000 + (c100 - c000) * tx, x10 = c010 + (c110 - c010) * tx;
    double x01 = c001 + (c101 - c001) * tx, x11 = c011 + (c111 - c011) * tx;
    double y0 = x00 + (x10 - x00) * ty, y1 = x01 + (x11 - x01) * ty;
    return y0 + (y1 - y0) * tz;
}
__host__ __device__ inline double texture(double x, double y, double z, double albedo, double wl, uint32_t seed) {
    double v = 0, amp = 1, f = 1.0 / wl, norm = 0;
    for (int o = 0; o < 5; o++) {
        v += amp * value_noise(x * f, y * f, z * f, seed + 101u * (uint32_t)o);
        norm += amp;
        amp *= 0.5; f *= 2.0;
    }
    double g = albedo + 0.45 * v / norm;
    return g < 0 ? 0 : (g > 1 ? 1 : g);
}

// nearest hit along o + t d (t > 0): returns false on a miss
__host__ __device__ inline bool cast(const Scene& s, const double* o, const double* d, double& tbest, double& tex) {
    tbest = 1e30;
    int kind = -1, idx = -1;
    for (int i = 0; i < 5; i++) {
        const Plane& p = s.planes[i];
        double den = p.n[0] * d[0] + p.n[1] * d[1] + p.n[2] * d[2];
        if (fabs(den) < 1e-12) continue;
        double t = (p.c - (p.n[0] * o[0] + p.n[1] * o[1] + p.n[2] * o[2])) / den;
        if (t > 1e-6 && t < tbest) { tbest = t; kind = 0; idx = i; }
    }
    for (int i = 0; i < s.nboxes; i++) {
        const Box& b = s.boxes[i];
        double lo[3], ld[3];
        double rel[3] = {o[0] - b.ctr[0], o[1] - b.ctr[1], o[2] - b.ctr[2]};
        for (int k = 0; k < 3; k++) {
            lo[k] = b.R[0 * 3 + k] * rel[0] + b.R[1 * 3 + k] * rel[1] + b.R[2 * 3 + k] * rel[2];
            ld[k] = b.R[0 * 3 + k] * d[0] + b.R[1 * 3 + k] * d[1] + b.R[2 * 3 + k] * d[2];
        }
        double t0 = -1e30, t1 = 1e30;
        bool miss = false;
        for (int k = 0; k < 3 && !miss; k++) {
            if (fabs(ld[k]) < 1e-12) {
                if (lo[k] < -b.h[k] || lo[k] > b.h[k]) miss = true;
                continue;
            }
            double a = (-b.h[k] - lo[k]) / ld[k], c = (b.h[k] - lo[k]) / ld[k];
            if (a > c) { double tmp = a; a = c; c = tmp; }
            t0 = t0 > a ? t0 : a;
            t1 = t1 < c ? t1 : c;
            if (t0 > t1) miss = true;
        }
        if (miss) continue;
        double t = t0 > 1e-6 ? t0 : t1;
        if (t > 1e-6 && t < tbest) { tbest = t; kind = 1; idx = i; }
    }
    if (kind < 0) return false;
    double P[3] = {o[0] + tbest * d[0], o[1] + tbest * d[1], o[2] + tbest * d[2]};
    if (kind == 0) {
        const Plane& p = s.planes[idx];
        tex = texture(P[0], P[1], P[2], p.albedo, p.wavelength, p.tseed);
    } else {
        const Box& b = s.boxes[idx];
        tex = texture(P[0] - b.ctr[0], P[1] - b.ctr[1], P[2] - b.ctr[2], b.albedo, b.wavelength, b.tseed);
    }
    return true;
}

// Counter-based standard normal sample for (seed, pixel): Box-Muller on two
// hashed uniforms.
__host__ __device__ inline double pixel_gauss(uint64_t seed, uint64_t pix) {
    Rng g(seed * 0x2545F4914F6CDD1Dull + pix * 0x9E3779B97F4A7C15ull + 12345ull);
    double u1 = g.uni(), u2 = g.uni();
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// One pixel of a frame: gray (2x2 supersampled) and depth in mm.
__host__ __device__ inline void render_pixel(const Scene& s, const double* T, int u, int v, double fx, double fy,
                                             double cx, double cy, uint64_t noise_seed, double noise_frac, int W,
                                             uint8_t& gray, uint16_t& depth) {
    double o[3] = {T[3], T[7], T[11]};
    double acc = 0;
    int hits = 0;
    for (int sy = 0; sy < 2; sy++)
        for (int sx = 0; sx < 2; sx++) {
            double pu = u - 0.25 + 0.5 * sx, pv = v - 0.25 + 0.5 * sy;
            double dc[3] = {(pu - cx) / fx, -(pv - cy) / fy, -1.0};
            double d[3];
            for (int k = 0; k < 3; k++) d[k] = T[k * 4 + 0] * dc[0] + T[k * 4 + 1] * dc[1] + T[k * 4 + 2] * dc[2];
            double t, tex;
            if (cast(s, o, d, t, tex)) { acc += tex; hits++; }
        }
    double gval = hits ? acc / hits : 0.0;
    gray = (uint8_t)floor(255.0 * gval + 0.5);
    double dc[3] = {(u - cx) / fx, -(v - cy) / fy, -1.0};
    double d[3];
    for (int k = 0; k < 3; k++) d[k] = T[k * 4 + 0] * dc[0] + T[k * 4 + 1] * dc[1] + T[k * 4 + 2] * dc[2];
    double t, tex;
    uint16_t dm = 0;
    if (cast(s, o, d, t, tex)) {
        double z = t;  // the camera-frame ray has unit -z component, so t is the depth
        if (noise_frac > 0) z += noise_frac * z * pixel_gauss(noise_seed, (uint64_t)v * (uint64_t)W + (uint64_t)u);
        if (z >= 500.0 && z <= 4000.0) dm = (uint16_t)floor(z + 0.5);
    }
    depth = dm;
}

}  // namespace synth
