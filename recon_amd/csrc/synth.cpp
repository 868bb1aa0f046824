// Seeded synthetic RGB-D scenes for the tracker's tests and benchmarks.
//
// A scene is a room of five slightly tilted planes (back wall, floor, ceiling,
// two side walls) with 2-5 randomly oriented boxes inside it. Surfaces carry a
// view-independent fractal value-noise texture evaluated at the 3D hit point,
// so brightness constancy holds between views. Frames are ray cast from a
// camera pose and returned as 8-bit gray (2x2 supersampled) and 16-bit depth in
// millimetres along the optical axis, valid in [500, 4000] mm, with i.i.d.
// Gaussian noise of a given fraction of depth (0 = no measurement).
//
// Camera frames follow the tracker's convention: x right, y up, the camera
// looks down -z. Poses are row-major 4x4 doubles mapping camera points to
// world points (T_wc).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct Rng {  // splitmix64 stream
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed ^ 0x9E3779B97F4A7C15ull) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double uni(double a, double b) { return a + (b - a) * uni(); }
    double gauss() {
        double u1 = uni(), u2 = uni();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

inline uint32_t hash3(int x, int y, int z, uint32_t seed) {
    uint32_t h = seed * 0x9E3779B1u;
    h ^= (uint32_t)x * 0x85EBCA77u; h = (h << 13) | (h >> 19); h *= 0xC2B2AE3Du;
    h ^= (uint32_t)y * 0x27D4EB2Fu; h = (h << 13) | (h >> 19); h *= 0x165667B1u;
    h ^= (uint32_t)z * 0x9E3779B1u; h = (h << 13) | (h >> 19); h *= 0x85EBCA77u;
    h ^= h >> 16; h *= 0x7FEB352Du; h ^= h >> 15; h *= 0x846CA68Bu; h ^= h >> 16;
    return h;
}
inline double lattice(int x, int y, int z, uint32_t seed) {
    return (double)(hash3(x, y, z, seed) >> 8) * (2.0 / 16777216.0) - 1.0;
}
inline double fade(double t) { return t * t * (3.0 - 2.0 * t); }
double value_noise(double x, double y, double z, uint32_t seed) {
    double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    double tx = fade(x - fx), ty = fade(y - fy), tz = fade(z - fz);
    double c[2][2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int d = 0; d < 2; d++) c[a][b][d] = lattice(ix + a, iy + b, iz + d, seed);
    double x00 = c[0][0][0] + (c[1][0][0] - c[0][0][0]) * tx;
    double x10 = c[0][1][0] + (c[1][1][0] - c[0][1][0]) * tx;
    double x01 = c[0][0][1] + (c[1][0][1] - c[0][0][1]) * tx;
    double x11 = c[0][1][1] + (c[1][1][1] - c[0][1][1]) * tx;
    double y0 = x00 + (x10 - x00) * ty, y1 = x01 + (x11 - x01) * ty;
    return y0 + (y1 - y0) * tz;
}

struct Plane { double n[3], c; double albedo, wavelength; uint32_t tseed; };
struct Box { double ctr[3], R[9], h[3]; double albedo, wavelength; uint32_t tseed; };

struct Scene {
    std::vector<Plane> planes;
    std::vector<Box> boxes;
};

void rot_axis_angle(const double ax[3], double ang, double R[9]) {
    double n = std::sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
    double x = ax[0] / n, y = ax[1] / n, z = ax[2] / n;
    double c = std::cos(ang), s = std::sin(ang), C = 1 - c;
    R[0] = c + x * x * C;     R[1] = x * y * C - z * s; R[2] = x * z * C + y * s;
    R[3] = y * x * C + z * s; R[4] = c + y * y * C;     R[5] = y * z * C - x * s;
    R[6] = z * x * C - y * s; R[7] = z * y * C + x * s; R[8] = c + z * z * C;
}
void random_unit(Rng& g, double v[3]) {
    double z = g.uni(-1, 1), a = g.uni(0, 6.283185307179586), r = std::sqrt(std::max(0.0, 1 - z * z));
    v[0] = r * std::cos(a); v[1] = r * std::sin(a); v[2] = z;
}

Scene make_scene(uint64_t seed) {
    Rng g(seed);
    Scene s;
    // room walls: inward normal n, plane n.x = c (c < 0 means the wall lies at distance -c)
    struct W { double n[3]; double dmin, dmax; };
    W walls[5] = {{{0, 0, 1}, 3000, 3800},    // back wall at z = -d
                  {{0, 1, 0}, 900, 1400},     // floor at y = -d
                  {{0, -1, 0}, 1200, 1800},   // ceiling at y = +d
                  {{1, 0, 0}, 1600, 2400},    // left wall at x = -d
                  {{-1, 0, 0}, 1600, 2400}};  // right wall at x = +d
    for (auto& w : walls) {
        double ax[3]; random_unit(g, ax);
        double R[9]; rot_axis_angle(ax, g.uni(0, 8.0) * M_PI / 180.0, R);
        Plane p;
        for (int i = 0; i < 3; i++) p.n[i] = R[i * 3 + 0] * w.n[0] + R[i * 3 + 1] * w.n[1] + R[i * 3 + 2] * w.n[2];
        p.c = -g.uni(w.dmin, w.dmax);
        p.albedo = g.uni(0.3, 0.7);
        p.wavelength = g.uni(150, 300);
        p.tseed = (uint32_t)g.next();
        s.planes.push_back(p);
    }
    int nb = 2 + (int)(g.next() % 4);
    for (int b = 0; b < nb; b++) {
        Box bx;
        double dist = g.uni(900, 2600);
        double azim = g.uni(-0.45, 0.45), elev = g.uni(-0.3, 0.3);
        bx.ctr[0] = dist * std::sin(azim);
        bx.ctr[1] = dist * std::sin(elev);
        bx.ctr[2] = -dist * std::cos(azim) * std::cos(elev);
        double ax[3]; random_unit(g, ax);
        rot_axis_angle(ax, g.uni(0, M_PI), bx.R);
        for (int i = 0; i < 3; i++) bx.h[i] = g.uni(100, 450);
        bx.albedo = g.uni(0.25, 0.75);
        bx.wavelength = g.uni(80, 200);
        bx.tseed = (uint32_t)g.next();
        s.boxes.push_back(bx);
    }
    return s;
}

inline double texture(double x, double y, double z, double albedo, double wl, uint32_t seed) {
    double v = 0, amp = 1, f = 1.0 / wl, norm = 0;
    for (int o = 0; o < 5; o++) {
        v += amp * value_noise(x * f, y * f, z * f, seed + 101u * (uint32_t)o);
        norm += amp;
        amp *= 0.5; f *= 2.0;
    }
    double g = albedo + 0.45 * v / norm;
    return g < 0 ? 0 : (g > 1 ? 1 : g);
}

// nearest hit along o + t d (t > 0); returns t and the texture value
bool cast(const Scene& s, const double o[3], const double d[3], double& tbest, double& tex) {
    tbest = 1e30;
    int kind = -1, idx = -1;
    for (size_t i = 0; i < s.planes.size(); i++) {
        const Plane& p = s.planes[i];
        double den = p.n[0] * d[0] + p.n[1] * d[1] + p.n[2] * d[2];
        if (std::fabs(den) < 1e-12) continue;
        double t = (p.c - (p.n[0] * o[0] + p.n[1] * o[1] + p.n[2] * o[2])) / den;
        if (t > 1e-6 && t < tbest) { tbest = t; kind = 0; idx = (int)i; }
    }
    for (size_t i = 0; i < s.boxes.size(); i++) {
        const Box& b = s.boxes[i];
        double lo[3], ld[3];
        double rel[3] = {o[0] - b.ctr[0], o[1] - b.ctr[1], o[2] - b.ctr[2]};
        for (int k = 0; k < 3; k++) {  // R^T
            lo[k] = b.R[0 * 3 + k] * rel[0] + b.R[1 * 3 + k] * rel[1] + b.R[2 * 3 + k] * rel[2];
            ld[k] = b.R[0 * 3 + k] * d[0] + b.R[1 * 3 + k] * d[1] + b.R[2 * 3 + k] * d[2];
        }
        double t0 = -1e30, t1 = 1e30;
        bool miss = false;
        for (int k = 0; k < 3 && !miss; k++) {
            if (std::fabs(ld[k]) < 1e-12) {
                if (lo[k] < -b.h[k] || lo[k] > b.h[k]) miss = true;
                continue;
            }
            double a = (-b.h[k] - lo[k]) / ld[k], c = (b.h[k] - lo[k]) / ld[k];
            if (a > c) std::swap(a, c);
            t0 = std::max(t0, a); t1 = std::min(t1, c);
            if (t0 > t1) miss = true;
        }
        if (miss) continue;
        double t = t0 > 1e-6 ? t0 : t1;
        if (t > 1e-6 && t < tbest) { tbest = t; kind = 1; idx = (int)i; }
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

void mat_exp_so3(const double w[3], double R[9]) {
    double th = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    if (th < 1e-15) { for (int i = 0; i < 9; i++) R[i] = (i % 4 == 0); return; }
    double ax[3] = {w[0] / th, w[1] / th, w[2] / th};
    rot_axis_angle(ax, th, R);
}
void compose(const double* A, const double* B, double* C) {
    double t[16];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            double s = 0;
            for (int k = 0; k < 4; k++) s += A[i * 4 + k] * B[k * 4 + j];
            t[i * 4 + j] = s;
        }
    std::memcpy(C, t, sizeof(t));
}

}  // namespace

extern "C" {

void* rs_scene_create(uint64_t seed) { return new Scene(make_scene(seed)); }
void rs_scene_destroy(void* s) { delete (Scene*)s; }

// Ray cast one frame. fx, fy, cx, cy: pinhole intrinsics in pixels (fx, fy > 0,
// row 0 at the top of the image); the pixel (u, v) looks along
// ((u-cx)/fx, -(v-cy)/fy, -1) in the camera frame.
void rs_render(const void* scene, const double* T_wc, int W, int H, double fx, double fy, double cx, double cy,
               uint64_t noise_seed, double noise_frac, uint8_t* gray, uint16_t* depth) {
    const Scene& s = *(const Scene*)scene;
    double o[3] = {T_wc[3], T_wc[7], T_wc[11]};
#pragma omp parallel for schedule(dynamic, 4)
    for (int v = 0; v < H; v++) {
        Rng g(noise_seed * 1000003ull + (uint64_t)v * 7919ull + 17ull);
        for (int u = 0; u < W; u++) {
            double acc = 0;
            int hits = 0;
            for (int sy = 0; sy < 2; sy++)
                for (int sx = 0; sx < 2; sx++) {
                    double pu = u - 0.25 + 0.5 * sx, pv = v - 0.25 + 0.5 * sy;
                    double dc[3] = {(pu - cx) / fx, -(pv - cy) / fy, -1.0};
                    double d[3];
                    for (int k = 0; k < 3; k++) d[k] = T_wc[k * 4 + 0] * dc[0] + T_wc[k * 4 + 1] * dc[1] + T_wc[k * 4 + 2] * dc[2];
                    double t, tex;
                    if (cast(s, o, d, t, tex)) { acc += tex; hits++; }
                }
            double gval = hits ? acc / hits : 0.0;
            gray[(size_t)v * W + u] = (uint8_t)std::lround(255.0 * gval);
            double dc[3] = {(u - cx) / fx, -(v - cy) / fy, -1.0};
            double d[3];
            for (int k = 0; k < 3; k++) d[k] = T_wc[k * 4 + 0] * dc[0] + T_wc[k * 4 + 1] * dc[1] + T_wc[k * 4 + 2] * dc[2];
            double t, tex;
            uint16_t dm = 0;
            if (cast(s, o, d, t, tex)) {
                double z = t;  // the camera-frame ray has unit -z component, so t is the depth
                if (noise_frac > 0) z += noise_frac * z * g.gauss();
                if (z >= 500.0 && z <= 4000.0) dm = (uint16_t)std::lround(z);
            }
            depth[(size_t)v * W + u] = dm;
        }
    }
}

// A rigid motion with axis uniform on the sphere, angle uniform in
// [0, theta_max] and translation uniform in the ball of radius t_max (mm).
void rs_sample_motion(uint64_t seed, double theta_max_deg, double t_max_mm, double* M) {
    Rng g(seed);
    double ax[3]; random_unit(g, ax);
    double ang = g.uni(0, theta_max_deg) * M_PI / 180.0;
    double R[9]; rot_axis_angle(ax, ang, R);
    double dir[3]; random_unit(g, dir);
    double r = t_max_mm * std::cbrt(g.uni());
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) M[i * 4 + j] = R[i * 3 + j];
        M[i * 4 + 3] = r * dir[i];
    }
    M[12] = M[13] = M[14] = 0; M[15] = 1;
}

// A smooth camera trajectory of n poses (T_wc), starting at the identity. Per
// frame the rotation increment stays below rot_max_deg and the translation
// increment below trans_max_mm; a weak spring keeps the camera near the origin
// looking down -z. seg_len > 0 alternates slow segments (limits scaled by
// slow_frac) with fast ones every seg_len frames.
void rs_trajectory(uint64_t seed, int n, double rot_max_deg, double trans_max_mm, int seg_len, double slow_frac,
                   double* poses) {
    Rng g(seed);
    double T[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    double w[3] = {0, 0, 0}, v[3] = {0, 0, 0};
    for (int k = 0; k < n; k++) {
        std::memcpy(poses + 16 * (size_t)k, T, sizeof(T));
        double scale = 1.0;
        if (seg_len > 0 && ((k / seg_len) % 2 == 0)) scale = slow_frac;
        double rmax = rot_max_deg * scale * M_PI / 180.0, tmax = trans_max_mm * scale;
        // the view direction's deviation from -z, and the position, pull back
        double fwd[3] = {-T[2], -T[6], -T[10]};
        for (int i = 0; i < 3; i++) {
            w[i] = 0.85 * w[i] + 0.35 * rmax * g.gauss();
            v[i] = 0.85 * v[i] + 0.35 * tmax * g.gauss() - 0.02 * T[i * 4 + 3];
        }
        // restoring rotation: axis fwd x (0,0,-1), expressed in the camera frame
        double want[3] = {0, 0, -1};
        double cr[3] = {fwd[1] * want[2] - fwd[2] * want[1], fwd[2] * want[0] - fwd[0] * want[2], fwd[0] * want[1] - fwd[1] * want[0]};
        for (int i = 0; i < 3; i++) {
            double c = T[0 * 4 + i] * cr[0] + T[1 * 4 + i] * cr[1] + T[2 * 4 + i] * cr[2];
            w[i] += 0.05 * c;
        }
        double wn = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
        if (wn > rmax) for (int i = 0; i < 3; i++) w[i] *= rmax / wn;
        double vn = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (vn > tmax) for (int i = 0; i < 3; i++) v[i] *= tmax / vn;
        double R[9]; mat_exp_so3(w, R);
        double D[16] = {R[0], R[1], R[2], 0, R[3], R[4], R[5], 0, R[6], R[7], R[8], 0, 0, 0, 0, 1};
        // translation step in world coordinates
        double Tn[16];
        compose(T, D, Tn);
        Tn[3] += v[0]; Tn[7] += v[1]; Tn[11] += v[2];
        std::memcpy(T, Tn, sizeof(T));
    }
}

}  // extern "C"
