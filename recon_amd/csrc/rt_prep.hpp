// Device bodies shared by the single-pair kernels (rt_kernels.hip) and the
// batched ones (rt_batch.hip): keyframe preparation blocks, the FP64 solve and
// the SE(3) update. Each body takes its block coordinates explicitly.
#pragma once
#include "rt_device.hpp"
#include "rt_kernels.hpp"

namespace rt {

template <int SRC>
__device__ __forceinline__ void convert_pyramid_block(int bx, int by, const uint8_t* __restrict__ g8,
                                                         const uint16_t* __restrict__ d16,
                                                         const float4* __restrict__ rgbd, int W, int H, int L,
                                                         const float* __restrict__ c, float* __restrict__ g0,
                                                         float* __restrict__ g1, float* __restrict__ g2,
                                                         float* __restrict__ g3, float* __restrict__ z,
                                                         const float* lut8 = nullptr) {
    // lut8 (SRC 3): (float)t / 255 for t = 0..255 in LDS, the IEEE quotients
    __shared__ float l2tile[16][17];
    int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    int x0 = (bx * 16 + tx) * 4, y0 = (by * 16 + ty) * 4;
    bool inside = x0 < W && y0 < H;
    float l2v = 0.0f;
    if (inside) {
        float v[4][4];
        float mn = c[MIND], mx = c[MAXD];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            int y = y0 + r;
            size_t o = (size_t)y * W + x0;
            if (SRC == 3) {  // 8-bit gray only: level 0 stays in its input form
                uchar4 gg = *reinterpret_cast<const uchar4*>(g8 + o);
                v[r][0] = lut8[gg.x]; v[r][1] = lut8[gg.y];
                v[r][2] = lut8[gg.z]; v[r][3] = lut8[gg.w];
            } else if (SRC == 0) {
                uchar4 gg = *reinterpret_cast<const uchar4*>(g8 + o);
                ushort4 dd = *reinterpret_cast<const ushort4*>(d16 + o);
                v[r][0] = u8_unit(gg.x); v[r][1] = u8_unit(gg.y);
                v[r][2] = u8_unit(gg.z); v[r][3] = u8_unit(gg.w);
                *reinterpret_cast<float4*>(g0 + o) = make_float4(v[r][0], v[r][1], v[r][2], v[r][3]);
                *reinterpret_cast<float4*>(z + o) =
                    make_float4(-(float)dd.x, -(float)dd.y, -(float)dd.z, -(float)dd.w);
            } else if (SRC == 1) {
                float zz[4];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    float4 p = rgbd[(size_t)(x0 + q) + (size_t)(H - 1 - y) * W];
                    v[r][q] = 0.3f * p.x + 0.59f * p.y + 0.11f * p.z;
                    zz[q] = -(p.w * (mx - mn) + mn);
                }
                *reinterpret_cast<float4*>(g0 + o) = make_float4(v[r][0], v[r][1], v[r][2], v[r][3]);
                *reinterpret_cast<float4*>(z + o) = make_float4(zz[0], zz[1], zz[2], zz[3]);
            } else {
                float4 gg = *reinterpret_cast<const float4*>(g0 + o);
                v[r][0] = gg.x; v[r][1] = gg.y; v[r][2] = gg.z; v[r][3] = gg.w;
            }
        }
        if (L > 1) {
            float l1[2][2];
            int W1 = W >> 1;
#pragma unroll
            for (int i = 0; i < 2; i++) {
#pragma unroll
                for (int j = 0; j < 2; j++)
                    l1[i][j] = (v[2 * i][2 * j] + v[2 * i][2 * j + 1] + v[2 * i + 1][2 * j] + v[2 * i + 1][2 * j + 1]) / 4.0f;
                *reinterpret_cast<float2*>(g1 + (size_t)(y0 / 2 + i) * W1 + x0 / 2) = make_float2(l1[i][0], l1[i][1]);
            }
            if (L > 2) {
                l2v = (l1[0][0] + l1[0][1] + l1[1][0] + l1[1][1]) / 4.0f;
                g2[(size_t)(y0 / 4) * (W >> 2) + x0 / 4] = l2v;
            }
        }
    }
    if (L > 3) {
        l2tile[ty][tx] = l2v;
        __syncthreads();
        if (inside && (tx & 1) == 0 && (ty & 1) == 0) {
            float l3 = (l2tile[ty][tx] + l2tile[ty][tx + 1] + l2tile[ty + 1][tx] + l2tile[ty + 1][tx + 1]) / 4.0f;
            g3[(size_t)(y0 / 8) * (W >> 3) + x0 / 8] = l3;
        }
    }
}


__device__ __forceinline__ void ref_features_block(int bx, int by, const float* __restrict__ z, const float* __restrict__ g0, int W,
                                                      int H, const float* __restrict__ c, float* __restrict__ gmag,
                                                      unsigned* __restrict__ ghist) {
    __shared__ unsigned h[256];
    h[threadIdx.x] = 0u;
    __syncthreads();
    int x = bx * 16 + (threadIdx.x & 15), y = by * 16 + (threadIdx.x >> 4);
    if (x < W && y < H) {
        PixFeat f = pixel_features(z, g0, W, H, x, y, c);
        gmag[x + (size_t)y * W] = f.gmag;
        atomicAdd(&h[grad_bin(f.gmag)], 1u);
    }
    __syncthreads();
    unsigned v = h[threadIdx.x];
    if (v) atomicAdd(&ghist[threadIdx.x], v);
}


// seekThreshold256Kernel2 (filterIndex256.cu:26-46): scanning the bins from the
// top, the threshold is the first bin whose running count exceeds N.
__device__ void find_threshold(const unsigned* __restrict__ ghist, int N, int& thr, int& excess) {
    __shared__ int sc[256];
    __shared__ int first;
    int t = threadIdx.x;
    if (t < 256) sc[t] = (int)ghist[255 - t];
    if (t == 0) first = 256;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        int v = (t < 256 && t >= off) ? sc[t - off] : 0;
        __syncthreads();
        if (t < 256) sc[t] += v;
        __syncthreads();
    }
    if (t < 256 && sc[t] > N && (t == 0 || sc[t - 1] <= N)) first = t;
    __syncthreads();
    if (first < 256) { thr = 255 - first; excess = sc[first] - N; }
    else { thr = 0; excess = sc[255] - N; }
    __syncthreads();
}

// filterIndexKernel4 (filterIndex256.cu:110-169), pass 1: per-segment counts of
// bins above and equal to the threshold. 512 segments of n/512 pixels.
__device__ __forceinline__ void select_count_block(int s, const float* __restrict__ gmag, int npix, const unsigned* ghist,
                                                      int N, int* seg) {
    int thr, excess;
    find_threshold(ghist, N, thr, excess);
    __shared__ int cA, cE;
    if (threadIdx.x == 0) { cA = 0; cE = 0; }
    __syncthreads();
    int P = npix / 512, b0 = s * P;
    int a = 0, e = 0;
    for (int i = b0 + threadIdx.x; i < b0 + P; i += 256) {
        unsigned bn = grad_bin(gmag[i]);
        a += bn > (unsigned)thr;
        e += bn == (unsigned)thr;
    }
    a = (int)wave_sum((float)a);
    e = (int)wave_sum((float)e);
    if (lane_id() == 0) { atomicAdd(&cA, a); atomicAdd(&cE, e); }
    __syncthreads();
    if (threadIdx.x == 0) {
        seg[s] = cA;
        seg[512 + s] = cE;
        if (s == 0) { seg[2560] = thr; seg[2561] = excess; }
    }
}

// pass 2 (one workgroup of 512): segment start offsets; the trim loop of
// filterIndex256.cu:141-152 run as written.
__device__ __forceinline__ void select_scan_block(int* seg) {
    __shared__ int a[512], m[512], sa[512], se[512];
    int t = threadIdx.x;
    a[t] = seg[t];
    m[t] = seg[512 + t];
    __syncthreads();
    if (t == 0) {
        int tooMuch = seg[2561];
        for (int i = 1; i < 512; i++) {
            if (tooMuch > 0 && m[i - 1] > 0) {
                m[i - 1] -= tooMuch;
                if (m[i - 1] < 0) { tooMuch += m[i - 1]; m[i - 1] = 0; }
                else tooMuch = 0;
            }
        }
    }
    sa[t] = a[t];
    se[t] = m[t];
    __syncthreads();
    for (int off = 1; off < 512; off <<= 1) {
        int va = t >= off ? sa[t - off] : 0, ve = t >= off ? se[t - off] : 0;
        __syncthreads();
        sa[t] += va; se[t] += ve;
        __syncthreads();
    }
    int totalA = sa[511];
    seg[1024 + t] = sa[t] - a[t];
    seg[1536 + t] = totalA + se[t] - m[t];
    seg[2048 + t] = m[t];
    if (t == 0) { seg[2562] = totalA; seg[2563] = totalA + se[511]; }
}

// pass 3: ordered scatter of each segment's above- and equal-bin indices.
__device__ __forceinline__ void select_scatter_block(int s, const float* __restrict__ gmag, int npix, int N,
                                                        const int* __restrict__ seg, int* __restrict__ sel) {
    __shared__ int wa[4], we[4];
    __shared__ int runA, runE;
    int t = threadIdx.x, wave = t >> 6, lane = lane_id();
    unsigned thr = (unsigned)seg[2560];
    int startA = seg[1024 + s], startE = seg[1536 + s], mE = seg[2048 + s];
    int P = npix / 512, b0 = s * P;
    if (t == 0) { runA = 0; runE = 0; }
    __syncthreads();
    uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    for (int base = b0; base < b0 + P; base += 256) {
        int i = base + t;
        bool in = i < b0 + P;
        unsigned bn = in ? grad_bin(gmag[i]) : 0u;
        bool isA = in && bn > thr, isE = in && bn == thr;
        uint64_t ma = __ballot(isA), me = __ballot(isE);
        if (lane == 0) { wa[wave] = __popcll(ma); we[wave] = __popcll(me); }
        __syncthreads();
        int pa = 0, pe = 0;
        for (int q = 0; q < wave; q++) { pa += wa[q]; pe += we[q]; }
        int rA = runA + pa + __popcll(ma & lt);
        int rE = runE + pe + __popcll(me & lt);
        if (isA) { int pos = startA + rA; if (pos < N) sel[pos] = i; }
        if (isE && rE < mE) { int pos = startE + rE; if (pos < N) sel[pos] = i; }
        __syncthreads();
        if (t == 0) { runA += wa[0] + wa[1] + wa[2] + wa[3]; runE += we[0] + we[1] + we[2] + we[3]; }
        __syncthreads();
    }
    if (s == 0) {
        int total = seg[2563];
        for (int k = (total > 0 ? total : 0) + t; k < N; k += 256) sel[k] = 0;
    }
}


template <class ZS, class GS>
__device__ __forceinline__ void ref_point(int i, const ZS& z, const GS& g0,
                                                    const float* __restrict__ g1, const float* __restrict__ g2,
                                                    const float* __restrict__ g3, int W, int H, int L, int N,
                                                    const float* __restrict__ c, float scale_in,
                                                    const int* __restrict__ sel, float4* __restrict__ pts,
                                                    float4* __restrict__ j0, float4* __restrict__ j1,
                                                    float4* __restrict__ j2, float4* __restrict__ j3) {
    int idx = ldg(sel, i);
    int xi = idx % W, yi = idx / W;
    PixFeat f = pixel_features(z, g0, W, H, xi, yi, c);
    float gx[4] = {f.gx, 0.0f, 0.0f, 0.0f}, gy[4] = {f.gy, 0.0f, 0.0f, 0.0f}, col[4] = {f.c1, 0.0f, 0.0f, 0.0f};
    const float* lv[4] = {nullptr, g1, g2, g3};
    {
        f3 r3 = mul4(c + TLR, f.p);
        f2 pu{r3.x / r3.z, r3.y / r3.z};
        f2 p1 = distort(pu, c + KCR, c + KR);
        if (L > 1 && in_bounds(p1, W, H)) {
            // Five bilinear samples (centre, N, W, E, S) per level with every tap
            // clamped into the image, because the reference's taps can leave the
            // image at the coarsest level near the border; together they read 12
            // distinct taps of the clamped 4x4 neighbourhood: load each once. The
            // taps of all levels are loaded before any is blended, so that their
            // loads are in flight together (levels past L read level 1 and are
            // discarded).
            f2 q = p1;
            float tp[3][12], fX[3], fY[3];
#pragma unroll
            for (int l = 1; l < 4; l++) {
                q.x = 0.5f * q.x + -0.25f;
                q.y = 0.5f * q.y + -0.25f;
                const int lq = l < L ? l : 1;
                int wl = W >> lq, hl = H >> lq;
                int xdi = (int)q.x, ydi = (int)q.y;
                fX[l - 1] = q.x - (float)xdi;
                fY[l - 1] = q.y - (float)ydi;
                const RT_GLOBAL float* im = gmem(lv[lq]);
                int xs[4], ys[4];
#pragma unroll
                for (int a = 0; a < 4; a++) {
                    xs[a] = min(max(xdi - 1 + a, 0), wl - 1);
                    ys[a] = min(max(ydi - 1 + a, 0), hl - 1) * wl;
                }
                float* t = tp[l - 1];
                t[0] = im[xs[1] + ys[0]]; t[1] = im[xs[2] + ys[0]];
                t[2] = im[xs[0] + ys[1]]; t[3] = im[xs[1] + ys[1]]; t[4] = im[xs[2] + ys[1]]; t[5] = im[xs[3] + ys[1]];
                t[6] = im[xs[0] + ys[2]]; t[7] = im[xs[1] + ys[2]]; t[8] = im[xs[2] + ys[2]]; t[9] = im[xs[3] + ys[2]];
                t[10] = im[xs[1] + ys[3]]; t[11] = im[xs[2] + ys[3]];
            }
            // every level is blended and kept by a select (levels past L stay 0),
            // so that no level's loads sink into a branch of their own
#pragma unroll
            for (int l = 1; l < 4; l++) {
                const float* t = tp[l - 1];
                const float a01 = t[0], a02 = t[1], a10 = t[2], a11 = t[3], a12 = t[4], a13 = t[5];
                const float a20 = t[6], a21 = t[7], a22 = t[8], a23 = t[9], a31 = t[10], a32 = t[11];
                const float fx = fX[l - 1], fy = fY[l - 1];
                float cc = blend(a11, a12, a21, a22, fx, fy);  // at (xdi, ydi)
                float cN = blend(a01, a02, a11, a12, fx, fy);  // (xdi, ydi - 1)
                float cW = blend(a10, a11, a20, a21, fx, fy);  // (xdi - 1, ydi)
                float cE = blend(a12, a13, a22, a23, fx, fy);  // (xdi + 1, ydi)
                float cS = blend(a21, a22, a31, a32, fx, fy);  // (xdi, ydi + 1)
                gx[l] = l < L ? (cE - cW) / 2.0f : 0.0f;
                gy[l] = l < L ? (cS - cN) / 2.0f : 0.0f;
                col[l] = l < L ? cc : 0.0f;
            }
        }
    }
    // Jacobian (warp.cu:502-645)
    const float* K = c + KR;
    const float* T = c + TLR;
    const float* kc = c + KCR;
    f3 r3 = mul4(T, f.p);
    r3.x *= scale_in; r3.y *= scale_in; r3.z *= scale_in;
    float dN0 = 1.0f / r3.z, dN2 = -r3.x / (r3.z * r3.z), dN4 = 1.0f / r3.z, dN5 = -r3.y / (r3.z * r3.z);
    float x = r3.x / r3.z, y = r3.y / r3.z;
    float x2 = x * x, y2 = y * y, x4 = x2 * x2, y4 = y2 * y2, r2 = x2 + y2, r4 = r2 * r2;
    float k0 = kc[0], k1 = kc[1], k4 = kc[4];
    float dD0 = 1.0f + k0 * (3.0f * x2 + y2) + k1 * (5.0f * x4 + 6.0f * x2 * y2 + y4) + k4 * r4 * (7.0f * x2 + y2);
    float dD1 = k0 * 2.0f * x * y + k1 * 4.0f * x * y * r2 + k4 * 6.0f * x * y * r4;
    float dD2 = k0 * 2.0f * y * x + k1 * 4.0f * x * y * r2 + k4 * 6.0f * x * y * r4;
    float dD3 = 1.0f + k0 * (3.0f * x2 + y2) + k1 * (5.0f * x4 + 6.0f * x2 * y2 + y4) + k4 * r4 * (7.0f * y2 + x2);
    f3 gens[6] = {{0.0f, -r3.z, r3.y}, {r3.z, 0.0f, -r3.x}, {-r3.y, r3.x, 0.0f},
                  {1.0f, 0.0f, 0.0f},  {0.0f, 1.0f, 0.0f},  {0.0f, 0.0f, 1.0f}};
    float J[4][6];
#pragma unroll
    for (int k = 0; k < 6; k++) {
        f3 dr3 = rot4(T, gens[k]);
        float px = dN0 * dr3.x + 0.0f * dr3.y + dN2 * dr3.z;
        float py = 0.0f * dr3.x + dN4 * dr3.y + dN5 * dr3.z;
        f3 dd2{dD0 * px + dD1 * py, dD2 * px + dD3 * py, 0.0f};
        f3 dp2 = mul3(K, dd2);
        J[0][k] = dp2.x * gx[0] + dp2.y * gy[0];
#pragma unroll
        for (int l = 1; l < 4; l++) {
            float sc = (float)(1 << l);
            J[l][k] = (dp2.x / sc) * gx[l] + (dp2.y / sc) * gy[l];
        }
    }
    stg(pts, i, make_float4(f.p.x, f.p.y, f.p.z, 0.0f));
    float4* jl[4] = {j0, j1, j2, j3};
    for (int l = 0; l < L; l++) {
        JacView v = jac_view(jl[l], N);
        stg(v.a, i, make_float4(J[l][0], J[l][1], J[l][2], J[l][3]));
        stg(v.b, i, make_float2(J[l][4], J[l][5]));
        stg(v.iref, i, col[l]);
    }
}


// ------------------------------------------------------------------ solve
// conjugateGradientKernel (warp.cu:926-975) and a textbook FP64 Cholesky.
// Every array the solve works on lives in a caller-provided workspace w (LDS
// in the kernels), so that one lane's serial FP64 solve does not hold ~200
// doubles in registers that the whole workgroup would then have to reserve.
// Workspace layout (doubles): see SOLVE_WS.
enum {
    WS_A = 0, WS_B = 36, WS_X = 42, WS_AG = 48, WS_BG = 84, WS_SOLVE = 90, WS_EXPM = 132, SOLVE_WS = 216
};
__device__ void solve_cg(const double* A, const double* bb, double* x, double* w) {
    double* r = w;
    double* dir = w + 6;
    double* Ad = w + 12;
    const double tol = 1e-8;
    for (int i = 0; i < 6; i++) { x[i] = 0.0; r[i] = bb[i]; dir[i] = bb[i]; }
    for (int step = 0; step < 6; step++) {
        for (int j = 0; j < 6; j++) {
            double s = 0.0;
            for (int k = 0; k < 6; k++) s += A[j * 6 + k] * dir[k];
            Ad[j] = s;
        }
        double rr = 0.0, Add = 0.0;
        for (int i = 0; i < 6; i++) rr += r[i] * r[i];
        for (int i = 0; i < 6; i++) Add += Ad[i] * dir[i];
        double dv = Add < 0 ? -Add : Add;
        double st = 0.0;
        if (dv > tol) st = rr / Add;
        for (int i = 0; i < 6; i++) r[i] -= st * Ad[i];
        double rr2 = 0.0;
        for (int i = 0; i < 6; i++) rr2 += r[i] * r[i];
        for (int i = 0; i < 6; i++) x[i] += st * dir[i];
        if (sqrt(rr2) < tol) return;
        double beta = rr2 / rr;
        for (int i = 0; i < 6; i++) dir[i] = r[i] + beta * dir[i];
    }
}
__device__ bool solve_cholesky(const double* A, const double* b, double* x, double* w) {
    double* Lm = w;
    double* y = w + 36;
    for (int i = 0; i < 36; i++) Lm[i] = 0.0;
    for (int j = 0; j < 6; j++) {
        double s = A[j * 6 + j];
        for (int k = 0; k < j; k++) s -= Lm[j * 6 + k] * Lm[j * 6 + k];
        if (!(s > 0.0)) { for (int i = 0; i < 6; i++) x[i] = 0.0; return false; }
        double dg = sqrt(s);
        Lm[j * 6 + j] = dg;
        for (int i = j + 1; i < 6; i++) {
            double t = A[i * 6 + j];
            for (int k = 0; k < j; k++) t -= Lm[i * 6 + k] * Lm[j * 6 + k];
            Lm[i * 6 + j] = t / dg;
        }
    }
    for (int i = 0; i < 6; i++) {
        double t = b[i];
        for (int k = 0; k < i; k++) t -= Lm[i * 6 + k] * y[k];
        y[i] = t / Lm[i * 6 + i];
    }
    for (int i = 5; i >= 0; i--) {
        double t = y[i];
        for (int k = i + 1; k < 6; k++) t -= Lm[k * 6 + i] * x[k];
        x[i] = t / Lm[i * 6 + i];
    }
    return true;
}
// normalizeMat4RealSafe (expmCuda.cu:1957-1980): column lengths taken before
// the cross products, with single-precision square roots.
__device__ void normalize_mat4_safe(double* T) {
    double u[3] = {T[0], T[4], T[8]}, v[3] = {T[1], T[5], T[9]}, w[3] = {T[2], T[6], T[10]};
    double lenU = (double)sqrtf((float)(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
    double lenW = (double)sqrtf((float)(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]));
    if (lenW > 1e-8f) { w[0] /= lenW; w[1] /= lenW; w[2] /= lenW; }
    else { w[0] = 0; w[1] = 0; w[2] = 1; }
    u[0] = v[1] * w[2] - v[2] * w[1];
    u[1] = -(v[0] * w[2] - v[2] * w[0]);
    u[2] = v[0] * w[1] - v[1] * w[0];
    if (lenU > 1e-8f) { u[0] /= lenU; u[1] /= lenU; u[2] /= lenU; }
    else { u[0] = 1; u[1] = 0; u[2] = 0; }
    v[0] = w[1] * u[2] - w[2] * u[1];
    v[1] = -(w[0] * u[2] - w[2] * u[0]);
    v[2] = w[0] * u[1] - w[1] * u[0];
    T[0] = u[0]; T[4] = u[1]; T[8] = u[2];
    T[1] = v[0]; T[5] = v[1]; T[9] = v[2];
    T[2] = w[0]; T[6] = w[1]; T[10] = w[2];
}
// Closed-form SE(3) exponential of generateA(x) (warp.cu:910-915); w: 36 doubles.
__device__ void se3_exp(const double* x, double* E, double* w) {
    double wx = x[0], wy = x[1], wz = x[2];
    double th2 = wx * wx + wy * wy + wz * wz, th = sqrt(th2);
    double a, b, cc;
    if (th < 1e-4) {
        a = 1.0 - th2 / 6.0 + th2 * th2 / 120.0;
        b = 0.5 - th2 / 24.0 + th2 * th2 / 720.0;
        cc = 1.0 / 6.0 - th2 / 120.0 + th2 * th2 / 5040.0;
    } else {
        double sn = sin(th), co = cos(th);
        a = sn / th; b = (1.0 - co) / th2; cc = (th - sn) / (th2 * th);
    }
    double* Wm = w;
    double* W2 = w + 9;
    double* Rm = w + 18;
    double* V = w + 27;
    Wm[0] = 0; Wm[1] = -wz; Wm[2] = wy; Wm[3] = wz; Wm[4] = 0; Wm[5] = -wx; Wm[6] = -wy; Wm[7] = wx; Wm[8] = 0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s += Wm[i * 3 + k] * Wm[k * 3 + j];
            W2[i * 3 + j] = s;
        }
    for (int i = 0; i < 9; i++) {
        double I = (i % 4 == 0) ? 1.0 : 0.0;
        Rm[i] = I + a * Wm[i] + b * W2[i];
        V[i] = I + b * Wm[i] + cc * W2[i];
    }
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) E[i * 4 + j] = Rm[i * 3 + j];
        E[i * 4 + 3] = V[i * 3 + 0] * x[3] + V[i * 3 + 1] * x[4] + V[i * 3 + 2] * x[5];
    }
    E[12] = 0; E[13] = 0; E[14] = 0; E[15] = 1;
}
// expmKernel (expmCuda.cu:1989-2048): T <- normalize(T * normalize(exp(A(x))) with t scaled);
// w: 84 doubles.
__device__ void expm_update(const double* x, float* TF, double scale, double* w) {
    double* T = w;
    double* Ti = w + 16;
    double* R = w + 32;
    for (int i = 0; i < 16; i++) T[i] = TF[i];
    double mx = 0;
    for (int i = 0; i < 6; i++) mx = fmax(mx, fabs(x[i]));
    if (mx != 0) {
        se3_exp(x, Ti, w + 48);
        normalize_mat4_safe(Ti);
        Ti[3] *= scale; Ti[7] *= scale; Ti[11] *= scale;
    } else {
        for (int i = 0; i < 16; i++) Ti[i] = (i % 5 == 0) ? 1.0 : 0.0;
    }
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            R[i * 4 + j] = T[i * 4 + 0] * Ti[0 * 4 + j] + T[i * 4 + 1] * Ti[1 * 4 + j] + T[i * 4 + 2] * Ti[2 * 4 + j] +
                           T[i * 4 + 3] * Ti[3 * 4 + j];
    normalize_mat4_safe(R);
    for (int i = 0; i < 16; i++) TF[i] = (float)R[i];
}
__device__ void unpack27(const double* s, double* A, double* b) {
    int k = 0;
    for (int r = 0; r < 6; r++)
        for (int c2 = r; c2 < 6; c2++) { A[r * 6 + c2] = s[k]; A[c2 * 6 + r] = s[k]; k++; }
    for (int r = 0; r < 6; r++) b[r] = s[21 + r];
}
// estimateIncrement (phototracker.cpp:646-664) with linearFuse (warp.cu:1498-1509);
// w: SOLVE_WS doubles. Returns false on a degenerate system (Cholesky).
__device__ bool solve_update(const double* sys, float* trel, const SolveCfg& cfg, int N1, int N2, double* w) {
    double* A = w + WS_A;
    double* b = w + WS_B;
    double* x = w + WS_X;
    unpack27(cfg.mode == 0 ? sys + 27 : sys, A, b);
    if (cfg.mode == 2) {
        double* Ag = w + WS_AG;
        double* bg = w + WS_BG;
        unpack27(sys + 27, Ag, bg);
        double s1 = (1.0 - (double)cfg.lambda) / (double)N1, s2 = (double)cfg.lambda / (double)N2;
        for (int k = 0; k < 36; k++) A[k] = s1 * A[k] + s2 * Ag[k];
        for (int k = 0; k < 6; k++) b[k] = s1 * b[k] + s2 * bg[k];
    }
    bool ok = true;
    if (cfg.solver == 1) solve_cg(A, b, x, w + WS_SOLVE);
    else ok = solve_cholesky(A, b, x, w + WS_SOLVE);
    expm_update(x, trel, (double)cfg.scale_out, w + WS_EXPM);
    return ok;
}

// expm_update with every matrix in registers instead of a workspace: the same
// operations in the same order (se3_exp, normalize_mat4_safe, the product
// T * Ti, normalize_mat4_safe). Run by every lane of a wave on the same x,
// so that no lane waits on the others; lane `writer` stores the pose.
__device__ __forceinline__ void normalize_mat4_safe_regs(double* T) {
    double u[3] = {T[0], T[4], T[8]}, v[3] = {T[1], T[5], T[9]}, w[3] = {T[2], T[6], T[10]};
    double lenU = (double)sqrtf((float)(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
    double lenW = (double)sqrtf((float)(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]));
    if (lenW > 1e-8f) { w[0] /= lenW; w[1] /= lenW; w[2] /= lenW; }
    else { w[0] = 0; w[1] = 0; w[2] = 1; }
    u[0] = v[1] * w[2] - v[2] * w[1];
    u[1] = -(v[0] * w[2] - v[2] * w[0]);
    u[2] = v[0] * w[1] - v[1] * w[0];
    if (lenU > 1e-8f) { u[0] /= lenU; u[1] /= lenU; u[2] /= lenU; }
    else { u[0] = 1; u[1] = 0; u[2] = 0; }
    v[0] = w[1] * u[2] - w[2] * u[1];
    v[1] = -(w[0] * u[2] - w[2] * u[0]);
    v[2] = w[0] * u[1] - w[1] * u[0];
    T[0] = u[0]; T[4] = u[1]; T[8] = u[2];
    T[1] = v[0]; T[5] = v[1]; T[9] = v[2];
    T[2] = w[0]; T[6] = w[1]; T[10] = w[2];
}
__device__ __forceinline__ void expm_update_regs(const double* x, float* TF, double scale, bool writer) {
    double Ti[16];
    double mx = 0;
#pragma unroll
    for (int i = 0; i < 6; i++) mx = fmax(mx, fabs(x[i]));
    if (mx != 0) {
        // se3_exp
        const double wx = x[0], wy = x[1], wz = x[2];
        const double th2 = wx * wx + wy * wy + wz * wz, th = sqrt(th2);
        double a, b, cc;
        if (th < 1e-4) {
            a = 1.0 - th2 / 6.0 + th2 * th2 / 120.0;
            b = 0.5 - th2 / 24.0 + th2 * th2 / 720.0;
            cc = 1.0 / 6.0 - th2 / 120.0 + th2 * th2 / 5040.0;
        } else {
            const double sn = sin(th), co = cos(th);
            a = sn / th; b = (1.0 - co) / th2; cc = (th - sn) / (th2 * th);
        }
        const double Wm[9] = {0, -wz, wy, wz, 0, -wx, -wy, wx, 0};
        double W2[9];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                double s2 = 0;
#pragma unroll
                for (int k = 0; k < 3; k++) s2 += Wm[i * 3 + k] * Wm[k * 3 + j];
                W2[i * 3 + j] = s2;
            }
        double V[9];
#pragma unroll
        for (int i = 0; i < 9; i++) {
            const double I = (i % 4 == 0) ? 1.0 : 0.0;
            Ti[(i / 3) * 4 + (i % 3)] = I + a * Wm[i] + b * W2[i];
            V[i] = I + b * Wm[i] + cc * W2[i];
        }
#pragma unroll
        for (int i = 0; i < 3; i++) Ti[i * 4 + 3] = V[i * 3 + 0] * x[3] + V[i * 3 + 1] * x[4] + V[i * 3 + 2] * x[5];
        Ti[12] = 0; Ti[13] = 0; Ti[14] = 0; Ti[15] = 1;
        normalize_mat4_safe_regs(Ti);
        Ti[3] *= scale; Ti[7] *= scale; Ti[11] *= scale;
    } else {
#pragma unroll
        for (int i = 0; i < 16; i++) Ti[i] = (i % 5 == 0) ? 1.0 : 0.0;
    }
    double R[16];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const double t0 = TF[i * 4 + 0], t1 = TF[i * 4 + 1], t2 = TF[i * 4 + 2], t3 = TF[i * 4 + 3];
#pragma unroll
        for (int j = 0; j < 4; j++) R[i * 4 + j] = t0 * Ti[0 * 4 + j] + t1 * Ti[1 * 4 + j] + t2 * Ti[2 * 4 + j] + t3 * Ti[3 * 4 + j];
    }
    normalize_mat4_safe_regs(R);
    if (writer) {
#pragma unroll
        for (int i = 0; i < 16; i++) TF[i] = (float)R[i];
    }
}

// One double of lane `src` (a compile-time constant after unrolling) to every lane.
__device__ __forceinline__ double lane_bcast(double v, int src) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
    return __hiloint2double(hi, lo);
}
// solve_update with the Cholesky solver, run by the 64 lanes of one wave: lane
// r < 6 holds row r of A and of L in registers, the factorisation computes a
// column's entries of all rows at once, and the substitutions exchange values
// by lane reads instead of a workspace. The system is assembled as in
// solve_update (photometric terms, geometric terms, or their linear fuse with
// the same operations), and every entry is computed by the same operations in
// the same order as in solve_cholesky, so the result is the same bit for bit;
// the exponential update runs in every lane's registers and lane 0 stores the
// pose. The CG solver runs solve_update on lane 0. The return value is
// meaningful in lane 0.
__device__ bool solve_update_wave(const double* sys, float* trel, const SolveCfg& cfg, int N1, int N2, double* w) {
    const int lane = lane_id();
    if (cfg.solver != 0) {
        bool ok = true;
        if (lane == 0) ok = solve_update(sys, trel, cfg, N1, N2, w);
        return ok;
    }
    const int r = lane < 6 ? lane : 5;
    // row r of A and b_r from the 27 packed terms (unpack27's order)
    const double* s0 = cfg.mode == 0 ? sys + 27 : sys;
    double Ar[6];
#pragma unroll
    for (int c = 0; c < 6; c++) {
        const int lo = r < c ? r : c, hi = r < c ? c : r;
        Ar[c] = s0[lo * 6 - lo * (lo - 1) / 2 + (hi - lo)];
    }
    double br = s0[21 + r];
    if (cfg.mode == 2) {
        // linearFuse (warp.cu:1498-1509) as in solve_update
        const double s1 = (1.0 - (double)cfg.lambda) / (double)N1, s2 = (double)cfg.lambda / (double)N2;
        const double* sg = sys + 27;
#pragma unroll
        for (int c = 0; c < 6; c++) {
            const int lo = r < c ? r : c, hi = r < c ? c : r;
            Ar[c] = s1 * Ar[c] + s2 * sg[lo * 6 - lo * (lo - 1) / 2 + (hi - lo)];
        }
        br = s1 * br + s2 * sg[21 + r];
    }
    double Lr[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    bool ok = true;
#pragma unroll
    for (int j = 0; j < 6; j++) {
        // lane j: s = A[j][j] - sum L[j][k]^2; lanes i > j: t = A[i][j] - sum L[i][k] L[j][k]
        double u = Ar[j];
#pragma unroll
        for (int k = 0; k < j; k++) u -= Lr[k] * lane_bcast(Lr[k], j);
        const double sj = lane_bcast(u, j);
        if (!(sj > 0.0)) { ok = false; break; }
        const double dg = sqrt(sj);
        const double q = u / dg;
        Lr[j] = r == j ? dg : (r > j ? q : Lr[j]);
    }
    double x[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (ok) {
        double yv = 0.0;
#pragma unroll
        for (int i = 0; i < 6; i++) {
            double t = br;
#pragma unroll
            for (int k = 0; k < i; k++) t -= Lr[k] * lane_bcast(yv, k);
            const double yi = t / Lr[i];
            yv = r == i ? yi : yv;
        }
#pragma unroll
        for (int i = 5; i >= 0; i--) {
            double t = lane_bcast(yv, i);
#pragma unroll
            for (int k = i + 1; k < 6; k++) t -= lane_bcast(Lr[i], k) * x[k];
            x[i] = t / lane_bcast(Lr[i], i);
        }
    }
    expm_update_regs(x, trel, (double)cfg.scale_out, lane == 0);
    return ok;
}

// ------------------------------------------------------------------ A10 + A11 per point
// interpolateResidualKernel2 (warp.cu:269-332) for one point: r = I_ref(l) -
// bilerp(I_cur(l)) (1.0 outside the view) and the zWeight from the nearest
// full-resolution current depth. a = 2^-l, b = (a-1)/2 (interpolateResidual2,
// warp.cu:1209-1211).
// SIMPLE: the calibration satisfies calib_simple, so r3 = T p, no distortion,
// and the depth-camera reprojection of the zWeight is the level-0 projection
// p2 itself.
template <bool SIMPLE = false, class GS, class ZS>
__device__ __forceinline__ void photo_point(float4 p4, float iref, const float* T, float a, float b, const GS& img,
                                            int wl, int hl, const ZS& zcur, int W, int H, const float* c, float& ir,
                                            float& zW) {
    f3 p3{p4.x, p4.y, p4.z};
    f3 r3ir = mul4(T, p3);
    f3 r3 = SIMPLE ? r3ir : mul4(c + TLR, r3ir);
    // the divisions as div_rn (the IEEE quotients; a zero or non-finite z gives
    // NaN where the division gives an infinity, which the finiteness test
    // below rejects alike)
    const float iz = 1.0f / r3.z;
    f2 pu{div_rn(r3.x, r3.z, iz), div_rn(r3.y, r3.z, iz)};
    f2 p2 = SIMPLE ? project_simple(pu, c + KR) : distort(pu, c + KCR, c + KR);
    f2 p{a * p2.x + b, a * p2.y + b};
    ir = 1.0f;
    zW = 0.0f;
    // a non-finite or huge projection is outside the view (the conversion
    // of such a value to int is undefined)
    bool fin = fabsf(p.x) < 1e9f && fabsf(p.y) < 1e9f;
    int xdi = fin ? (int)p.x : -1, ydi = fin ? (int)p.y : -1;
    if (xdi >= 0 && ydi >= 0 && xdi < wl - 1 && ydi < hl - 1) {
        float fx = p.x - (float)xdi, fy = p.y - (float)ydi;
        float color = bilerp(img, wl, xdi, ydi, fx, fy);
        ir = iref - color;
        f3 pl2;
        if (SIMPLE) {
            pl2 = f3{p2.x, p2.y, 1.0f};
        } else {
            f3 rl3 = mul4(c + TRL, r3);
            const float il = 1.0f / rl3.z;
            rl3.x = div_rn(rl3.x, rl3.z, il); rl3.y = div_rn(rl3.y, rl3.z, il); rl3.z = 1.0f;
            pl2 = mul3(c + KL, rl3);
        }
        bool fin2 = fabsf(pl2.x) < 1e9f && fabsf(pl2.y) < 1e9f;
        int xdi2 = fin2 ? (int)(pl2.x + 0.5f) : -1, ydi2 = fin2 ? (int)(pl2.y + 0.5f) : -1;
        if (xdi2 >= 0 && ydi2 >= 0 && xdi2 < W && ydi2 < H) {
            float zc = depth_tap(zcur, xdi2, ydi2, W);
            float zerr = zc - r3.z;
            zerr *= zerr;
            if (zerr < 10000.0f) zW = 1.0f - div_rn(zerr, 10000.0f, 1.0f / 10000.0f);
        }
    }
}
// MAD scale (histogram64.cu:82): 1.4826 (4m+3)/255 at 64 bins; (m+0.75)/255
// with 256 bins.
__device__ __forceinline__ float mad_scale(int m, int bins) {
    return bins == 64 ? 1.4826f * (float)(4 * m + 3) / 255.0f : 1.4826f * ((float)m + 0.75f) / 255.0f;
}
// generateWeights64Kernel (histogram64.cu:79-94): Tukey (1-(u/4.6851)^2) zW
// inside the threshold, or Huber sqrt(min(1, 1.345/u)) zW.
// The divisions by s (> 0: the MAD scale is at least 1.4826 * 3 / 255) and by
// the Tukey constant run as div_rn (the same quotients); 1 / s is the same for
// every point of an iteration.
__device__ __forceinline__ float robust_weight(float r, float zw, float s, int robust) {
    float u = div_rn(fabsf(r), s, 1.0f / s);
    float w = 0.0f;
    if (robust == 0) {
        const float bb = 4.6851f;
        if (u <= bb) {
            float ub = div_rn(u, bb, 1.0f / bb), ub2 = ub * ub;
            w = (1.0f - ub2) * zw;
        }
    } else {
        float k = 1.345f;
        float hw = (u <= k) ? 1.0f : k / u;
        w = sqrtf(hw) * zw;
    }
    return w;
}


// ------------------------------------------------------------------ A18-A23 per pixel / point
// convertZMapToXYZ6Kernel (rgb2gray.cu:595-624) + computeNormalsCudaKernel
// (d2z.cu:503-537) for pixel (x, y): position (x, y, z, mask) and normal from
// the depth source zs (z < 0 valid).
template <class ZS>
__device__ __forceinline__ void ref_cloud_pixel(const ZS& zs, int W, int H, int x, int y, const float* __restrict__ c,
                                                float4* __restrict__ pos, float4* __restrict__ nrm) {
    size_t o = (size_t)x + (size_t)y * W;
    f3 p;
    bool valid = backproject(zs[o], x, y, c, p);
    float mask = valid ? 1.0f : 0.0f;
    float4 n = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (x < 1 || y < 1 || x > W - 2 || y > H - 2) {
        mask = 0.0f;
    } else {
        f3 a, b2, cc, dd;
        backproject(zs[o - 1], x - 1, y, c, a);
        backproject(zs[o + 1], x + 1, y, c, b2);
        backproject(zs[o - W], x, y - 1, c, cc);
        backproject(zs[o + W], x, y + 1, c, dd);
        if (a.z < 0.0f && b2.z < 0.0f && cc.z < 0.0f && dd.z < 0.0f) {
            f3 u{b2.x - a.x, b2.y - a.y, b2.z - a.z};
            f3 v{dd.x - cc.x, dd.y - cc.y, dd.z - cc.z};
            float nx = u.y * v.z - u.z * v.y;
            float ny = -(u.x * v.z - u.z * v.x);
            float nz = u.x * v.y - u.y * v.x;
            float len = nx * nx + ny * ny + nz * nz;
            len = 1.0f / sqrtf(len);
            n = make_float4(nx * len, ny * len, nz * len, 0.0f);
        } else {
            mask = 0.0f;
        }
    }
    pos[o] = make_float4(p.x, p.y, p.z, mask);
    nrm[o] = n;
}
// genPointsCloud6Kernel (d2z.cu:436-469) for pixel i: every z < 0 is valid.
template <class ZS>
__device__ __forceinline__ void cur_cloud_pixel(const ZS& zs, int W, int i, const float* __restrict__ c,
                                                float4* __restrict__ pos) {
    int x = i % W, y = i / W;
    float zz = zs[(size_t)i];
    if (zz < 0.0f) {
        float fx = c[KL], fy = c[KL + 4], cx = c[KL + 2], cy = c[KL + 5];
        pos[i] = make_float4(((float)x - cx) * zz / fx, ((float)y - cy) * zz / fy, zz, 1.0f);
    } else {
        pos[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}
// downSampleCloud6Kernel (d2z.cu:557-588) for destination pixel i of a
// Wd x Hd level: masked 2x2 mean, normal renormalised (sn / dn may be null).
__device__ __forceinline__ void cloud_down_pixel(const float4* __restrict__ sp, const float4* __restrict__ sn, int Wd,
                                                 int i, float4* __restrict__ dp, float4* __restrict__ dn) {
    int x = i % Wd, y = i / Wd;
    int Ws = Wd * 2;
    size_t h = (size_t)x * 2 + (size_t)y * 2 * Ws;
    size_t hs[4] = {h, h + 1, h + Ws, h + Ws + 1};
    float d[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    float n = 0.0f;
#pragma unroll
    for (int t = 0; t < 4; t++) {
        float4 p = sp[hs[t]];
        if (p.w != 0.0f) {
            n++;
            d[0] += p.x; d[1] += p.y; d[2] += p.z;
            if (sn) { float4 q = sn[hs[t]]; d[3] += q.x; d[4] += q.y; d[5] += q.z; }
        }
    }
    if (n > 0.0f) {
        for (int k = 0; k < 6; k++) d[k] /= n;
        dp[i] = make_float4(d[0], d[1], d[2], 1.0f);
        if (dn) {
            float len = d[3] * d[3] + d[4] * d[4] + d[5] * d[5];
            len = 1.0f / sqrtf(len);
            dn[i] = make_float4(d[3] * len, d[4] * len, d[5] * len, 0.0f);
        }
    } else {
        dp[i] = make_float4(d[0], d[1], d[2], 0.0f);
        if (dn) dn[i] = make_float4(d[3], d[4], d[5], 0.0f);
    }
}
// icpResidualKernel2 (rgb2gray.cu:223-260) for keyframe pixel o of a w x h
// level (a = 2^-l, b = (a - 1) / 2): valid when the keyframe point exists and
// the four current taps it projects to are valid; r = (bilerp(c) - T p) . (R n).
__device__ __forceinline__ bool icp_mask_pixel(const float4* __restrict__ rp, const float4* __restrict__ rn,
                                               const float4* __restrict__ cp, int w, int h, int o, const float* T,
                                               const float* c, float a, float b, float& r) {
    float4 p = rp[o];
    if (p.w == 0.0f) return false;
    float4 nn = rn[o];
    f3 p3{p.x, p.y, p.z}, n3{nn.x, nn.y, nn.z};
    f3 q3 = mul4(T, p3);
    f3 cn = rot4(T, n3);
    const float* K = c + KL;
    float qx = q3.x / q3.z, qy = q3.y / q3.z;
    float px = a * (K[0] * qx + K[2]) + b;
    float py = a * (K[4] * qy + K[5]) + b;
    if (!(px >= 0.0f && px < (float)(w - 1) && py >= 0.0f && py < (float)(h - 1))) return false;
    int xdi = (int)px, ydi = (int)py;
    int o2 = xdi + ydi * w;
    float4 c0 = cp[o2], c1 = cp[o2 + 1], c2 = cp[o2 + w], c3 = cp[o2 + w + 1];
    if (!(c0.w != 0.0f && c1.w != 0.0f && c2.w != 0.0f && c3.w != 0.0f)) return false;
    float fx = px - (float)xdi, fy = py - (float)ydi;
    f3 cc{blend(c0.x, c1.x, c2.x, c3.x, fx, fy), blend(c0.y, c1.y, c2.y, c3.y, fx, fy),
          blend(c0.z, c1.z, c2.z, c3.z, fx, fy)};
    f3 dd{cc.x - q3.x, cc.y - q3.y, cc.z - q3.z};
    r = dd.x * cn.x + dd.y * cn.y + dd.z * cn.z;
    return true;
}
// icpResidualIndexKernel2 (rgb2gray.cu:323-390) for keyframe pixel o with its
// squared mask residual x2, the final Student-t weight (clamped to 1) from the
// variance sum v3 over n points: scaled residual and the 6 Jacobian entries
// (zeros when the point leaves the view, its first tap is invalid, the weight
// vanishes or |dz| >= 300 mm).
__device__ __forceinline__ void icp_point_terms(const float4* __restrict__ rp, const float4* __restrict__ rn,
                                                const float4* __restrict__ cp, int w, int h, int o, float x2, float v3,
                                                int n, const float* T, const float* c, float a, float b, float s,
                                                float& res, float* J) {
    float4 p = rp[o], nn = rn[o];
    f3 p3{p.x, p.y, p.z}, rnv{nn.x, nn.y, nn.z};
    f3 q3 = mul4(T, p3);
    f3 cn = rot4(T, rnv);
    const float* K = c + KL;
    float qx = q3.x / q3.z, qy = q3.y / q3.z;
    float px = a * (K[0] * qx + K[2]) + b;
    float py = a * (K[4] * qy + K[5]) + b;
    res = 0.0f;
#pragma unroll
    for (int k = 0; k < 6; k++) J[k] = 0.0f;
    if (px >= 0.0f && px < (float)(w - 1) && py >= 0.0f && py < (float)(h - 1)) {
        float wt = fminf(6 / (5 + x2 / (v3 / (float)n)), 1.0f);
        int xdi = (int)px, ydi = (int)py;
        float fx = px - (float)xdi, fy = py - (float)ydi;
        int o2 = xdi + ydi * w;
        float4 c0 = cp[o2], c1 = cp[o2 + 1], c2 = cp[o2 + w], c3 = cp[o2 + w + 1];
        f3 cc{blend(c0.x, c1.x, c2.x, c3.x, fx, fy), blend(c0.y, c1.y, c2.y, c3.y, fx, fy),
              blend(c0.z, c1.z, c2.z, c3.z, fx, fy)};
        if (c0.w != 0.0f && wt > 0.0f && fabsf(cc.z - q3.z) < 300.0f) {
            f3 d{cc.x - q3.x, cc.y - q3.y, cc.z - q3.z};
            res = s * wt * (d.x * cn.x + d.y * cn.y + d.z * cn.z);
            f3 gp[3] = {{0.0f, -p3.z, p3.y}, {p3.z, 0.0f, -p3.x}, {-p3.y, p3.x, 0.0f}};
            f3 gn[3] = {{0.0f, -rnv.z, rnv.y}, {rnv.z, 0.0f, -rnv.x}, {-rnv.y, rnv.x, 0.0f}};
#pragma unroll
            for (int k = 0; k < 3; k++) {
                f3 dB = rot4(T, gp[k]);
                f3 dN = rot4(T, gn[k]);
                J[k] = -s * wt * (-cn.x * dB.x - cn.y * dB.y - cn.z * dB.z + d.x * dN.x + d.y * dN.y + d.z * dN.z);
            }
            f3 e[3] = {{1.0f, 0.0f, 0.0f}, {0.0f, 1.0f, 0.0f}, {0.0f, 0.0f, 1.0f}};
#pragma unroll
            for (int k = 0; k < 3; k++) {
                f3 dB = rot4(T, e[k]);
                J[3 + k] = -wt * (-cn.x * dB.x - cn.y * dB.y - cn.z * dB.z);
            }
        }
    }
}

}  // namespace rt
