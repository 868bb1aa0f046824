// Seeded synthetic RGB-D frames (scene model in synth_scene.hpp): scene and
// motion sampling on the host, rendering on the host (threads) or the GPU.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>
#include <thread>
#include <vector>
#include "synth_scene.hpp"

using namespace synth;

namespace {

void rot_axis_angle(const double ax[3], double ang, double R[9]) {
    double n = sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
    double x = ax[0] / n, y = ax[1] / n, z = ax[2] / n;
    double c = cos(ang), s = sin(ang), C = 1 - c;
    R[0] = c + x * x * C;     R[1] = x * y * C - z * s; R[2] = x * z * C + y * s;
    R[3] = y * x * C + z * s; R[4] = c + y * y * C;     R[5] = y * z * C - x * s;
    R[6] = z * x * C - y * s; R[7] = z * y * C + x * s; R[8] = c + z * z * C;
}
void random_unit(Rng& g, double v[3]) {
    double z = g.uni(-1, 1), a = g.uni(0, 6.283185307179586), r = sqrt(fmax(0.0, 1 - z * z));
    v[0] = r * cos(a); v[1] = r * sin(a); v[2] = z;
}
double gauss(Rng& g) {
    double u1 = g.uni(), u2 = g.uni();
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

Scene make_scene(uint64_t seed) {
    Rng g(seed);
    Scene s;
    memset(&s, 0, sizeof(s));
    struct Wl { double n[3]; double dmin, dmax; };
    const Wl walls[5] = {{{0, 0, 1}, 3000, 3800},    // back wall at z = -d
                         {{0, 1, 0}, 900, 1400},     // floor at y = -d
                         {{0, -1, 0}, 1200, 1800},   // ceiling at y = +d
                         {{1, 0, 0}, 1600, 2400},    // left wall at x = -d
                         {{-1, 0, 0}, 1600, 2400}};  // right wall at x = +d
    for (int w = 0; w < 5; w++) {
        double ax[3]; random_unit(g, ax);
        double R[9]; rot_axis_angle(ax, g.uni(0, 8.0) * M_PI / 180.0, R);
        Plane& p = s.planes[w];
        for (int i = 0; i < 3; i++)
            p.n[i] = R[i * 3 + 0] * walls[w].n[0] + R[i * 3 + 1] * walls[w].n[1] + R[i * 3 + 2] * walls[w].n[2];
        p.c = -g.uni(walls[w].dmin, walls[w].dmax);
        p.albedo = g.uni(0.3, 0.7);
        p.wavelength = g.uni(150, 300);
        p.tseed = (uint32_t)g.next();
    }
    s.nboxes = 2 + (int)(g.next() % 4);
    for (int b = 0; b < s.nboxes; b++) {
        Box& bx = s.boxes[b];
        double dist = g.uni(900, 2600);
        double azim = g.uni(-0.45, 0.45), elev = g.uni(-0.3, 0.3);
        bx.ctr[0] = dist * sin(azim);
        bx.ctr[1] = dist * sin(elev);
        bx.ctr[2] = -dist * cos(azim) * cos(elev);
        double ax[3]; random_unit(g, ax);
        rot_axis_angle(ax, g.uni(0, M_PI), bx.R);
        for (int i = 0; i < 3; i++) bx.h[i] = g.uni(100, 450);
        bx.albedo = g.uni(0.25, 0.75);
        bx.wavelength = g.uni(80, 200);
        bx.tseed = (uint32_t)g.next();
    }
    return s;
}

void mat_exp_so3(const double w[3], double R[9]) {
    double th = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
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
    memcpy(C, t, sizeof(t));
}

struct FrameArgs {
    double T[16];
    double fx, fy, cx, cy, noise;
    uint64_t seed;
    int W, H;
};

__global__ __launch_bounds__(256) void k_render(Scene s, FrameArgs a, uint8_t* gray, uint16_t* depth) {
    int u = blockIdx.x * 16 + (threadIdx.x & 15), v = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (u >= a.W || v >= a.H) return;
    uint8_t g; uint16_t d;
    render_pixel(s, a.T, u, v, a.fx, a.fy, a.cx, a.cy, a.seed, a.noise, a.W, g, d);
    gray[(size_t)v * a.W + u] = g;
    depth[(size_t)v * a.W + u] = d;
}

}  // namespace

extern "C" {

void* rs_scene_create(uint64_t seed) { return new Scene(make_scene(seed)); }
void rs_scene_destroy(void* s) { delete (Scene*)s; }

// Ray cast one frame on the host. fx, fy, cx, cy: pinhole intrinsics in
// pixels (fx, fy > 0, row 0 at the top); pixel (u, v) looks along
// ((u-cx)/fx, -(v-cy)/fy, -1) in the camera frame.
void rs_render(const void* scene, const double* T_wc, int W, int H, double fx, double fy, double cx, double cy,
               uint64_t noise_seed, double noise_frac, uint8_t* gray, uint16_t* depth) {
    const Scene& s = *(const Scene*)scene;
    unsigned nt = std::thread::hardware_concurrency();
    if (nt < 1) nt = 1;
    if (nt > 16) nt = 16;
    std::vector<std::thread> th;
    for (unsigned k = 0; k < nt; k++)
        th.emplace_back([&, k]() {
            for (int v = (int)k; v < H; v += (int)nt)
                for (int u = 0; u < W; u++)
                    render_pixel(s, T_wc, u, v, fx, fy, cx, cy, noise_seed, noise_frac, W, gray[(size_t)v * W + u],
                                 depth[(size_t)v * W + u]);
        });
    for (auto& t : th) t.join();
}

// The same frame rendered by the GPU into device buffers (asynchronous on
// stream). Returns 0 or a HIP error code.
int rs_render_device(const void* scene, const double* T_wc, int W, int H, double fx, double fy, double cx, double cy,
                     uint64_t noise_seed, double noise_frac, uint8_t* gray_dev, uint16_t* depth_dev, void* stream) {
    FrameArgs a;
    memcpy(a.T, T_wc, sizeof(a.T));
    a.fx = fx; a.fy = fy; a.cx = cx; a.cy = cy; a.noise = noise_frac; a.seed = noise_seed; a.W = W; a.H = H;
    dim3 grid((W + 15) / 16, (H + 15) / 16);
    k_render<<<grid, 256, 0, (hipStream_t)stream>>>(*(const Scene*)scene, a, gray_dev, depth_dev);
    return (int)hipGetLastError();
}

// A rigid motion with axis uniform on the sphere, angle uniform in
// [0, theta_max] and translation uniform in the ball of radius t_max (mm).
void rs_sample_motion(uint64_t seed, double theta_max_deg, double t_max_mm, double* M) {
    Rng g(seed);
    double ax[3]; random_unit(g, ax);
    double ang = g.uni(0, theta_max_deg) * M_PI / 180.0;
    double R[9]; rot_axis_angle(ax, ang, R);
    double dir[3]; random_unit(g, dir);
    double r = t_max_mm * cbrt(g.uni());
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) M[i * 4 + j] = R[i * 3 + j];
        M[i * 4 + 3] = r * dir[i];
    }
    M[12] = M[13] = M[14] = 0; M[15] = 1;
}

// A smooth camera trajectory of n poses (T_wc) from the identity. Per frame
// the rotation increment stays below rot_max_deg and the translation increment
// below trans_max_mm; a weak spring keeps the camera near the origin looking
// down -z. seg_len > 0 alternates slow segments (limits scaled by slow_frac)
// with fast ones every seg_len frames.
void rs_trajectory(uint64_t seed, int n, double rot_max_deg, double trans_max_mm, int seg_len, double slow_frac,
                   double* poses) {
    Rng g(seed);
    double T[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    double w[3] = {0, 0, 0}, v[3] = {0, 0, 0};
    for (int k = 0; k < n; k++) {
        memcpy(poses + 16 * (size_t)k, T, sizeof(T));
        double scale = 1.0;
        if (seg_len > 0 && ((k / seg_len) % 2 == 0)) scale = slow_frac;
        double rmax = rot_max_deg * scale * M_PI / 180.0, tmax = trans_max_mm * scale;
        double fwd[3] = {-T[2], -T[6], -T[10]};
        for (int i = 0; i < 3; i++) {
            w[i] = 0.85 * w[i] + 0.35 * rmax * gauss(g);
            v[i] = 0.85 * v[i] + 0.35 * tmax * gauss(g) - 0.02 * T[i * 4 + 3];
        }
        double want[3] = {0, 0, -1};
        double cr[3] = {fwd[1] * want[2] - fwd[2] * want[1], fwd[2] * want[0] - fwd[0] * want[2],
                        fwd[0] * want[1] - fwd[1] * want[0]};
        for (int i = 0; i < 3; i++) w[i] += 0.05 * (T[0 * 4 + i] * cr[0] + T[1 * 4 + i] * cr[1] + T[2 * 4 + i] * cr[2]);
        double wn = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
        if (wn > rmax) for (int i = 0; i < 3; i++) w[i] *= rmax / wn;
        double vn = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (vn > tmax) for (int i = 0; i < 3; i++) v[i] *= tmax / vn;
        double R[9]; mat_exp_so3(w, R);
        double D[16] = {R[0], R[1], R[2], 0, R[3], R[4], R[5], 0, R[6], R[7], R[8], 0, 0, 0, 0, 1};
        double Tn[16];
        compose(T, D, Tn);
        Tn[3] += v[0]; Tn[7] += v[1]; Tn[11] += v[2];
        memcpy(T, Tn, sizeof(T));
    }
}

}  // extern "C"
