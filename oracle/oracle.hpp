// CPU oracle for the recon keyframe-tracking hot path.
//
// TEST INFRASTRUCTURE ONLY. This is the checker that the HIP implementation is
// compared against; nothing in the product library links or calls it. Only the
// test-suite, the smoke check and the CPU-baseline leg of the benchmark may load
// it (through liboracle.so).
//
// Every function restates one reference kernel or host routine of
// ttykkala/recon and cites the file:line it follows. The functions are
// templated on the arithmetic type R:
//   R = float  : a literal FP32 transcription. Expressions keep the reference's
//                evaluation order and are compiled with -ffp-contract=off, so a
//                device kernel written in the same order reproduces them bit for
//                bit (the HIP library is compiled with the same flag).
//   R = double : the FP64 oracle used to state the error of every FP32 result.
// Integer / index work (histograms, point selection, compaction) is exact in both.
//
// Conventions are the reference's own: OpenGL camera frame (camera looks down
// -z, depth in millimetres, z < 0 for valid points), intrinsics stored in the
// 96-float calibration block of tracker/calib/calib.h:20-36 (fx < 0, fy > 0).
//
// Where the reference reads outside an image (attribute taps at the coarsest
// level for points near the border, tracker/kernels/d2z.cu:905-934), the
// oracle clamps the tap to the image; that edge is "parity unpinned" against the
// reference, which reads whatever memory is adjacent.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace oracle {

// Offsets of the calibration block (tracker/calib/calib.h:20-36).
enum {
    KR_OFFSET = 0, iKR_OFFSET = 9, KcR_OFFSET = 18, TLR_OFFSET = 26, C1_OFFSET = 42,
    KL_OFFSET = 43, iKL_OFFSET = 52, C0_OFFSET = 61, MIND_OFFSET = 62, MAXD_OFFSET = 63,
    TRL_OFFSET = 64, CALIB_SIZE = 96
};

// Vertex record slots (tracker/rendering/VertexBuffer2.h:23).
enum {
    VB_X = 0, VB_Y = 1, VB_Z = 2, VB_P2X = 6, VB_P2Y = 7, VB_GX1 = 11, VB_GY1 = 12, VB_GMAG = 13,
    VB_C1 = 14, VB_GX2 = 15, VB_GY2 = 16, VB_C2 = 17, VB_GX3 = 18, VB_GY3 = 19, VB_C3 = 20,
    VB_GX4 = 21, VB_GY4 = 22, VB_C4 = 23, VB_STRIDE = 24
};

template <class R> struct V3 { R x, y, z; };
template <class R> struct V2 { R x, y; };

// ---------------------------------------------------------------- helpers
// matrixMultVec4 / matrixRot4 / matrixMultVec3 (tracker/kernels/kernelUtils.h:232-253)
template <class R, class M>
inline V3<R> mul4(const M* T, V3<R> v) {
    V3<R> r;
    r.x = (R)T[0] * v.x + (R)T[1] * v.y + (R)T[2] * v.z + (R)T[3];
    r.y = (R)T[4] * v.x + (R)T[5] * v.y + (R)T[6] * v.z + (R)T[7];
    r.z = (R)T[8] * v.x + (R)T[9] * v.y + (R)T[10] * v.z + (R)T[11];
    return r;
}
template <class R, class M>
inline V3<R> rot4(const M* T, V3<R> v) {
    V3<R> r;
    r.x = (R)T[0] * v.x + (R)T[1] * v.y + (R)T[2] * v.z;
    r.y = (R)T[4] * v.x + (R)T[5] * v.y + (R)T[6] * v.z;
    r.z = (R)T[8] * v.x + (R)T[9] * v.y + (R)T[10] * v.z;
    return r;
}
template <class R, class M>
inline V3<R> mul3(const M* K, V3<R> v) {
    V3<R> r;
    r.x = (R)K[0] * v.x + (R)K[1] * v.y + (R)K[2] * v.z;
    r.y = (R)K[3] * v.x + (R)K[4] * v.y + (R)K[5] * v.z;
    r.z = (R)K[6] * v.x + (R)K[7] * v.y + (R)K[8] * v.z;
    return r;
}

// distortPoint (tracker/kernels/kernelUtils.h:394-408)
template <class R>
inline V2<R> distort(V2<R> pu, const float* kc, const float* K) {
    R dx = pu.x * pu.x, dy = pu.y * pu.y;
    R r2 = dx + dy, r4 = r2 * r2, r6 = r4 * r2;
    R rd = (R)1 + (R)kc[0] * r2 + (R)kc[1] * r4 + (R)kc[4] * r6;
    V2<R> pd;
    pd.x = (R)K[0] * pu.x * rd + (R)K[2];
    pd.y = (R)K[4] * pu.y * rd + (R)K[5];
    return pd;
}

// inBounds (tracker/kernels/kernelUtils.h:341-344)
template <class R>
inline bool in_bounds(V2<R> p, int w, int h) {
    return p.x >= (R)10 && p.x < (R)(w - 10) && p.y >= (R)10 && p.y < (R)(h - 10);
}

// bilinearInterpolation(xdi,ydi,fx,fy,...) (tracker/kernels/kernelUtils.h:267-272)
template <class R, class P>
inline R bilerp(const P* img, int width, int xdi, int ydi, R fx, R fy) {
    int o = xdi + ydi * width;
    R v0 = (R)img[o], v1 = (R)img[o + 1], v2 = (R)img[o + width], v3 = (R)img[o + width + 1];
    return ((R)1 - fx) * ((R)1 - fy) * v0 + fx * ((R)1 - fy) * v1 + ((R)1 - fx) * fy * v2 + fx * fy * v3;
}
// Same, with each tap clamped into [0,w-1]x[0,h-1] (used only where the
// reference would read outside the image).
template <class R, class P>
inline R bilerp_clamped(const P* img, int w, int h, int xdi, int ydi, R fx, R fy) {
    auto at = [&](int x, int y) -> R {
        x = std::min(std::max(x, 0), w - 1);
        y = std::min(std::max(y, 0), h - 1);
        return (R)img[x + y * w];
    };
    R v0 = at(xdi, ydi), v1 = at(xdi + 1, ydi), v2 = at(xdi, ydi + 1), v3 = at(xdi + 1, ydi + 1);
    return ((R)1 - fx) * ((R)1 - fy) * v0 + fx * ((R)1 - fy) * v1 + ((R)1 - fx) * fy * v2 + fx * fy * v3;
}

// ---------------------------------------------------------------- A1
// rgbTex2GrayHdrZKernel (tracker/kernels/rgb2gray.cu:64-76): float4 RGBD texture,
// read Y-flipped (row H-1-y), gray = 0.3r+0.59g+0.11b, z = -(w(max-min)+min).
template <class R>
void convert_rgbd(const float* rgbd, int W, int H, const float* calib, R* gray, R* z) {
    R mn = (R)calib[MIND_OFFSET], mx = (R)calib[MAXD_OFFSET];
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const float* p = rgbd + 4 * (size_t)(x + (H - 1 - y) * W);
            size_t o = (size_t)x + (size_t)y * W;
            gray[o] = (R)0.3f * (R)p[0] + (R)0.59f * (R)p[1] + (R)0.11f * (R)p[2];
            z[o] = -((R)p[3] * (mx - mn) + mn);
        }
}
// The same conversion for the boundary's u8 gray / u16 depth (mm) input, which
// is already top-down: gray = g/255 (the [0,1] range of the texture path),
// z = -depth (0 = no measurement, rejected downstream exactly as the texture
// path's w = 0 encoding is).
template <class R>
void convert_u8u16(const uint8_t* g, const uint16_t* d, int n, R* gray, R* z) {
    for (int i = 0; i < n; i++) {
        gray[i] = (R)g[i] / (R)255.0f;
        z[i] = -(R)d[i];
    }
}

// ---------------------------------------------------------------- A2
// fastDownSampleHdrKernel (tracker/kernels/downSample2Cuda.cu:35-42)
template <class R>
void downsample2(const R* hi, int Wh, R* lo, int Wl, int Hl) {
    for (int y = 0; y < Hl; y++)
        for (int x = 0; x < Wl; x++) {
            size_t h = (size_t)(2 * x) + (size_t)(2 * y) * Wh;
            lo[x + (size_t)y * Wl] = (hi[h] + hi[h + 1] + hi[h + Wh] + hi[h + Wh + 1]) / (R)4.0f;
        }
}

// ---------------------------------------------------------------- A3
// convertZMapToXYZKernel (tracker/kernels/rgb2gray.cu:570-593)
template <class R>
inline void backproject(R z, int xi, int yi, const float* calib, R out[4]) {
    R mn = (R)calib[MIND_OFFSET], mx = (R)calib[MAXD_OFFSET];
    if (z < -mn && z > -mx) {
        R fx = (R)calib[KL_OFFSET], fy = (R)calib[KL_OFFSET + 4];
        R cx = (R)calib[KL_OFFSET + 2], cy = (R)calib[KL_OFFSET + 5];
        out[0] = ((R)xi - cx) * z / fx;
        out[1] = ((R)yi - cy) * z / fy;
        out[2] = z;
        out[3] = 1;
    } else {
        out[0] = out[1] = out[2] = out[3] = 0;
    }
}

// ---------------------------------------------------------------- A4
// xyz2CloudKernel (tracker/kernels/d2z.cu:235-331): per-pixel vertex record with
// level-0 gray, central-difference gradient and gradient magnitude.
// rec: W*H*24 values (the reference's 24-slot record layout).
template <class R>
void vertex_records(const R* z, const R* gray0, int W, int H, const float* calib, R* rec) {
    const float* T = calib + TLR_OFFSET;
    const float* KR = calib + KR_OFFSET;
    const float* kc = calib + KcR_OFFSET;
    R maxD = (R)calib[MAXD_OFFSET];
    for (int yi = 0; yi < H; yi++)
        for (int xi = 0; xi < W; xi++) {
            size_t off = (size_t)xi + (size_t)yi * W;
            R* v = rec + off * VB_STRIDE;
            R p4[4];
            backproject<R>(z[off], xi, yi, calib, p4);
            V3<R> p3{p4[0], p4[1], p4[2]};
            if (p3.z < 0) {
                V3<R> r3 = mul4<R>(T, p3);
                V2<R> pu{r3.x / r3.z, r3.y / r3.z};
                V2<R> p = distort<R>(pu, kc, KR);
                if (in_bounds<R>(p, W, H)) {
                    int xdi = (int)p.x, ydi = (int)p.y;
                    R fx = p.x - (R)xdi, fy = p.y - (R)ydi;
                    R c1 = bilerp<R>(gray0, W, xdi, ydi, fx, fy);
                    R cN = bilerp<R>(gray0, W, (int)p.x, (int)(p.y - (R)1.0f), fx, fy);
                    R cE = bilerp<R>(gray0, W, (int)(p.x + (R)1.0f), (int)p.y, fx, fy);
                    R cW = bilerp<R>(gray0, W, (int)(p.x - (R)1.0f), (int)p.y, fx, fy);
                    R cS = bilerp<R>(gray0, W, (int)p.x, (int)(p.y + (R)1.0f), fx, fy);
                    R gx = (cE - cW) / (R)2.0f;
                    R gy = (cS - cN) / (R)2.0f;
                    for (int i = 0; i < VB_STRIDE; i++) v[i] = 0;
                    v[VB_X] = p3.x; v[VB_Y] = p3.y; v[VB_Z] = p3.z;
                    v[VB_P2X] = p.x; v[VB_P2Y] = p.y;
                    v[VB_GX1] = gx; v[VB_GY1] = gy;
                    v[VB_GMAG] = std::min(std::fabs(gy) + std::fabs(gx), (R)1.0f);
                    v[VB_C1] = c1;
                    continue;
                }
            }
            for (int i = 0; i < VB_STRIDE; i++) v[i] = 0;
            v[VB_Z] = -maxD;
        }
}

// ---------------------------------------------------------------- A6
// Gradient-magnitude histogram (tracker/kernels/cuda_hist.cu:100-152,
// gpu_hist.cu:113-160): bin = (uint)(g*(bins-1)+0.5f). The multiply-add is
// taken as one fused operation, as the reference's compiler contracts it.
inline unsigned grad_bin(float g, int nbins) {
    return (unsigned)std::fmaf(g, (float)(nbins - 1), 0.5f);
}
inline unsigned grad_bin(double g, int nbins) {
    return (unsigned)(g * (double)(nbins - 1) + 0.5);
}
template <class R>
inline void grad_histogram(const R* g, int n, int nbins, uint32_t* hist) {
    std::memset(hist, 0, sizeof(uint32_t) * nbins);
    for (int i = 0; i < n; i++) {
        unsigned b = grad_bin(g[i], nbins);
        if (b < (unsigned)nbins) hist[b]++;
    }
}

// ---------------------------------------------------------------- A7
// seekThreshold256Kernel2 + filterIndexKernel4 (tracker/kernels/filterIndex256.cu:26-46,
// 110-169, host wrapper :171-194). 512 segments of n/512 pixels (integer
// division: a tail of n%512 pixels is never considered); indices whose bin is
// above the threshold bin come first (ascending), then threshold-bin indices of
// each segment, after the trim loop of :141-152 removed part of the excess from
// the earliest segments (its running excess becomes the trimmed segment's own
// count when a segment holds fewer than the excess, and the last segment is
// never trimmed). The list is cut to nselect entries (the reference sets the
// element count to nselect). Returns threshold bin and excess through the
// out-parameters.
template <class R>
inline void select_points(const R* g, int n, const uint32_t* hist, int nbins, int nselect,
                          int* out, int* thr_out = nullptr, int* excess_out = nullptr) {
    unsigned thr = 0;
    int acc = 0;
    for (int i = nbins - 1; i >= 0; i--) {
        acc += (int)hist[i];
        if (acc > nselect) { thr = (unsigned)i; break; }
    }
    int excess = acc - nselect;
    if (thr_out) *thr_out = (int)thr;
    if (excess_out) *excess_out = excess;
    const int S = 512;
    int P = n / S;
    std::vector<int> cnt(S), mc(S);
    std::vector<std::vector<int>> above(S), eq(S);
    for (int s = 0; s < S; s++) {
        for (int i = s * P; i < s * P + P; i++) {
            unsigned b = grad_bin(g[i], nbins);
            if (b > thr) above[s].push_back(i);
            else if (b == thr) eq[s].push_back(i);
        }
        cnt[s] = (int)above[s].size();
        mc[s] = (int)eq[s].size();
    }
    int tooMuch = excess;
    for (int i = 1; i < S; i++) {
        if (tooMuch > 0 && mc[i - 1] > 0) {
            mc[i - 1] -= tooMuch;
            if (mc[i - 1] < 0) { tooMuch += mc[i - 1]; mc[i - 1] = 0; }
            else tooMuch = 0;
        }
    }
    int k = 0;
    for (int s = 0; s < S && k < nselect; s++)
        for (int j = 0; j < cnt[s] && k < nselect; j++) out[k++] = above[s][j];
    for (int s = 0; s < S && k < nselect; s++)
        for (int j = 0; j < mc[s] && k < nselect; j++) out[k++] = eq[s][j];
    for (; k < nselect; k++) out[k] = 0;
}

// ---------------------------------------------------------------- A8
// addVertexAttributesWithoutNormalsKernel3/4 (tracker/kernels/d2z.cu:700-800, 805-936):
// gray and gradient at levels 1..L-1 for the selected points, p_{l+1} = 0.5 p_l - 0.25.
template <class R>
void point_attributes(const int* idx, int npts, R* rec, const R* const* gray, int W, int H, int L,
                      const float* calib) {
    const float* T = calib + TLR_OFFSET;
    const float* KR = calib + KR_OFFSET;
    const float* kc = calib + KcR_OFFSET;
    for (int k = 0; k < npts; k++) {
        R* v = rec + (size_t)idx[k] * VB_STRIDE;
        V3<R> p3{v[VB_X], v[VB_Y], v[VB_Z]};
        V3<R> r3 = mul4<R>(T, p3);
        V2<R> pu{r3.x / r3.z, r3.y / r3.z};
        V2<R> p = distort<R>(pu, kc, KR);
        R out[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        if (in_bounds<R>(p, W, H)) {
            V2<R> q = p;
            for (int l = 1; l < L; l++) {
                q.x = (R)0.5f * q.x + (R)-0.25f;
                q.y = (R)0.5f * q.y + (R)-0.25f;
                int wl = W >> l, hl = H >> l;
                int xdi = (int)q.x, ydi = (int)q.y;
                R fX = q.x - (R)xdi, fY = q.y - (R)ydi;
                const R* img = gray[l];
                R c = bilerp_clamped<R>(img, wl, hl, xdi, ydi, fX, fY);
                R cN = bilerp_clamped<R>(img, wl, hl, xdi, ydi - 1, fX, fY);
                R cW = bilerp_clamped<R>(img, wl, hl, xdi - 1, ydi, fX, fY);
                R cE = bilerp_clamped<R>(img, wl, hl, xdi + 1, ydi, fX, fY);
                R cS = bilerp_clamped<R>(img, wl, hl, xdi, ydi + 1, fX, fY);
                out[l - 1][0] = (cE - cW) / (R)2.0f;
                out[l - 1][1] = (cS - cN) / (R)2.0f;
                out[l - 1][2] = c;
            }
        }
        for (int l = 1; l < 4; l++) {
            v[VB_GX2 + 3 * (l - 1)] = out[l - 1][0];
            v[VB_GY2 + 3 * (l - 1)] = out[l - 1][1];
            v[VB_C2 + 3 * (l - 1)] = out[l - 1][2];
        }
    }
}

// ---------------------------------------------------------------- A9
// precomputeJacobian4Kernel (tracker/kernels/warp.cu:502-645): J_l[k][i], SoA of
// 6 rows x npts per level; level l divides the projected derivative by 2^l.
template <class R>
void jacobians(const int* idx, int npts, const R* rec, const float* calib, float optScaleIn, int L,
               R* const* J) {
    const float* K = calib + KR_OFFSET;
    const float* T = calib + TLR_OFFSET;
    const float* kc = calib + KcR_OFFSET;
    for (int i = 0; i < npts; i++) {
        const R* v = rec + (size_t)idx[i] * VB_STRIDE;
        V3<R> p3{v[VB_X], v[VB_Y], v[VB_Z]};
        V3<R> r3 = mul4<R>(T, p3);
        r3.x *= (R)optScaleIn; r3.y *= (R)optScaleIn; r3.z *= (R)optScaleIn;
        R gX[4] = {v[VB_GX1], v[VB_GX2], v[VB_GX3], v[VB_GX4]};
        R gY[4] = {v[VB_GY1], v[VB_GY2], v[VB_GY3], v[VB_GY4]};
        R dN[6];
        dN[0] = (R)1.0f / r3.z; dN[1] = 0; dN[2] = -r3.x / (r3.z * r3.z);
        dN[3] = 0; dN[4] = (R)1.0f / r3.z; dN[5] = -r3.y / (r3.z * r3.z);
        R x = r3.x / r3.z, y = r3.y / r3.z;
        R x2 = x * x, y2 = y * y, x4 = x2 * x2, y4 = y2 * y2, r2 = x2 + y2, r4 = r2 * r2;
        R k0 = (R)kc[0], k1 = (R)kc[1], k4 = (R)kc[4];
        R dD[4];
        dD[0] = (R)1 + k0 * ((R)3 * x2 + y2) + k1 * ((R)5 * x4 + (R)6 * x2 * y2 + y4) + k4 * r4 * ((R)7 * x2 + y2);
        dD[1] = k0 * (R)2 * x * y + k1 * (R)4 * x * y * r2 + k4 * (R)6 * x * y * r4;
        dD[2] = k0 * (R)2 * y * x + k1 * (R)4 * x * y * r2 + k4 * (R)6 * x * y * r4;
        dD[3] = (R)1 + k0 * ((R)3 * x2 + y2) + k1 * ((R)5 * x4 + (R)6 * x2 * y2 + y4) + k4 * r4 * ((R)7 * y2 + x2);
        V3<R> gens[6] = {{0, -r3.z, r3.y}, {r3.z, 0, -r3.x}, {-r3.y, r3.x, 0},
                         {1, 0, 0},        {0, 1, 0},        {0, 0, 1}};
        for (int k = 0; k < 6; k++) {
            V3<R> dr3 = rot4<R>(T, gens[k]);
            V3<R> dpn2, dd2;
            dpn2.x = dN[0] * dr3.x + dN[1] * dr3.y + dN[2] * dr3.z;
            dpn2.y = dN[3] * dr3.x + dN[4] * dr3.y + dN[5] * dr3.z;
            dpn2.z = 0;
            dd2.x = dD[0] * dpn2.x + dD[1] * dpn2.y;
            dd2.y = dD[2] * dpn2.x + dD[3] * dpn2.y;
            dd2.z = 0;
            V3<R> dp2 = mul3<R>(K, dd2);
            for (int l = 0; l < L; l++) {
                R val;
                if (l == 0) val = dp2.x * gX[0] + dp2.y * gY[0];
                else {
                    R s = (R)(float)(1 << l);
                    val = (dp2.x / s) * gX[l] + (dp2.y / s) * gY[l];
                }
                J[l][(size_t)k * npts + i] = val;
            }
        }
    }
}

// ---------------------------------------------------------------- A10
// interpolateResidualKernel2 (tracker/kernels/warp.cu:269-332) with the level
// mapping of interpolateResidual2 (:1160-1218): r = I_ref(l) - bilerp(I_cur(l)),
// 1.0 outside the field of view; zWeight from the nearest full-resolution current
// depth in the IR view.
template <class R, class M>
void photo_residual(const int* idx, int npts, const R* rec, const M* Tm, const float* calib, int layer,
                    const R* img, int wl, int hl, const R* zcur, int W, int H, R* res, R* zw) {
    const float* KR = calib + KR_OFFSET;
    const float* kc = calib + KcR_OFFSET;
    const float* TLR = calib + TLR_OFFSET;
    const float* KL = calib + KL_OFFSET;
    const float* TRL = calib + TRL_OFFSET;
    int coff = VB_C1 + 3 * layer;
    float af = 1.0f / (float)(1 << layer);
    float bf = 0.5f * (af - 1.0f);
    R a = (R)af, b = (R)bf;
    for (int i = 0; i < npts; i++) {
        const R* v = rec + (size_t)idx[i] * VB_STRIDE;
        V3<R> p3{v[VB_X], v[VB_Y], v[VB_Z]};
        V3<R> r3ir = mul4<R>(Tm, p3);
        V3<R> r3 = mul4<R>(TLR, r3ir);
        V2<R> pu{r3.x / r3.z, r3.y / r3.z};
        V2<R> p2 = distort<R>(pu, kc, KR);
        V2<R> p{a * p2.x + b, a * p2.y + b};
        R ir = (R)1.0f;
        // a non-finite or huge projection is outside the view (its conversion
        // to int is undefined)
        bool fin = std::fabs(p.x) < (R)1e9f && std::fabs(p.y) < (R)1e9f;
        int xdi = fin ? (int)p.x : -1, ydi = fin ? (int)p.y : -1;
        R zW = 0;
        if (xdi >= 0 && ydi >= 0 && xdi < wl - 1 && ydi < hl - 1) {
            R fx = p.x - (R)xdi, fy = p.y - (R)ydi;
            R color = bilerp<R>(img, wl, xdi, ydi, fx, fy);
            ir = v[coff] - color;
            V3<R> rl3 = mul4<R>(TRL, r3);
            rl3.x /= rl3.z; rl3.y /= rl3.z; rl3.z = 1;
            V3<R> pl2 = mul3<R>(KL, rl3);
            bool fin2 = std::fabs(pl2.x) < (R)1e9f && std::fabs(pl2.y) < (R)1e9f;
            int xdi2 = fin2 ? (int)(pl2.x + (R)0.5f) : -1, ydi2 = fin2 ? (int)(pl2.y + (R)0.5f) : -1;
            if (xdi2 >= 0 && ydi2 >= 0 && xdi2 < W && ydi2 < H) {
                R zc = zcur[xdi2 + (size_t)ydi2 * W];
                R zerr = zc - r3.z;
                zerr *= zerr;
                if (zerr < (R)(100 * 100)) zW = (R)1.0f - zerr / ((R)100.0f * (R)100.0f);
            }
        }
        res[i] = ir;
        zw[i] = zW;
    }
}

// ---------------------------------------------------------------- A11
// generateWeights64 (tracker/kernels/histogram64.cu:41-94, 115-158):
// bin = (int)((bins-1)|r|), median bin m = first bin whose running count exceeds
// n/2, s = 1.4826 (4m+3)/255 at 64 bins (the reference); the 256-bin variant
// uses s = 1.4826 (m+0.75)/255, the same 3/4-of-bin position on the 255 scale.
// Tukey: w = (1-(u/4.6851)^2) zW for u = |r|/s <= 4.6851 (weight then multiplies
// both J and r, so the IRLS weight is its square, Q2). Huber: w = sqrt(hw) zW
// with hw = min(1, 1.345/u), so the effective IRLS weight is Huber's.
enum { ROBUST_TUKEY = 0, ROBUST_HUBER = 1 };
template <class R>
inline int mad_median_bin(const R* res, int n, int bins, uint32_t* hist_out = nullptr) {
    std::vector<uint32_t> h(bins, 0);
    for (int i = 0; i < n; i++) {
        int b = (int)((R)(float)(bins - 1) * std::fabs(res[i]));
        if (b >= 0 && b < bins) h[b]++;
    }
    int m = 0, acc = 0;
    for (int i = 0; i < bins; i++) {
        acc += (int)h[i];
        if (acc > n / 2) { m = i; break; }
    }
    if (hist_out) std::memcpy(hist_out, h.data(), sizeof(uint32_t) * bins);
    return m;
}
template <class R>
inline R mad_scale(int m, int bins) {
    if (bins == 64) return (R)1.4826f * (R)(float)(4 * m + 3) / (R)255.0f;
    return (R)1.4826f * ((R)(float)m + (R)0.75f) / (R)255.0f;
}
template <class R>
void robust_weights(const R* res, const R* zw, int n, int m, int bins, int robust, R* w, R* wr) {
    R s = mad_scale<R>(m, bins);
    for (int i = 0; i < n; i++) {
        R r = res[i];
        R u = std::fabs(r) / s;
        R wi = 0;
        if (robust == ROBUST_TUKEY) {
            R bb = (R)4.6851f;
            if (u <= bb) {
                R ub = u / bb, ub2 = ub * ub;
                wi = ((R)1.0f - ub2) * zw[i];
            }
        } else {
            R k = (R)1.345f;
            R hw = (u <= k) ? (R)1.0f : k / u;
            wi = std::sqrt(hw) * zw[i];
        }
        wr[i] = r * wi;
        w[i] = wi;
    }
}

// ---------------------------------------------------------------- A12/A13
// weightJacobianKernel (tracker/kernels/warp.cu:770-780) and the 27 dot products
// of JTJCuda / JTresidualCuda (:1453-1487): A = sum Jw Jw^T (symmetric), b = sum
// Jw * wr with Jw = J*w. Per-point products are formed in R and accumulated in
// double (the order-independent FP64 reference for the FP32 device sums).
template <class R>
void normal_equations(const R* const* Jrows, const R* w, const R* wr, int n, double A[36], double b[6]) {
    for (int k = 0; k < 36; k++) A[k] = 0;
    for (int k = 0; k < 6; k++) b[k] = 0;
    for (int i = 0; i < n; i++) {
        R jw[6];
        for (int k = 0; k < 6; k++) jw[k] = Jrows[k][i] * w[i];
        for (int r = 0; r < 6; r++) {
            for (int c = r; c < 6; c++) A[r * 6 + c] += (double)(jw[r] * jw[c]);
            b[r] += (double)(jw[r] * wr[i]);
        }
    }
    for (int r = 0; r < 6; r++)
        for (int c = 0; c < r; c++) A[r * 6 + c] = A[c * 6 + r];
}

// ---------------------------------------------------------------- A14
// linearFuseKernel / linearFuseCuda (tracker/kernels/warp.cu:918-924,1498-1509;
// shapetracker.cpp:392-394): A = (1-l)/N1 A_photo + l/N2 A_icp, same for b, with
// N1 the configured photometric point count (Q10).
inline void linear_fuse(double A1[36], double b1[6], int N1, const double A2[36], const double b2[6], int N2,
                        double lambda) {
    double s1 = (1.0 - lambda) / (double)N1, s2 = lambda / (double)N2;
    for (int k = 0; k < 36; k++) A1[k] = s1 * A1[k] + s2 * A2[k];
    for (int k = 0; k < 6; k++) b1[k] = s1 * b1[k] + s2 * b2[k];
}

// ---------------------------------------------------------------- A15
// conjugateGradientKernel (tracker/kernels/warp.cu:926-975): 6-step FP64 CG,
// x0 = 0, tol 1e-8 (:1536-1537).
inline void solve_cg(const double A[36], const double bb[6], double x[6], double tol = 1e-8, int maxSteps = 6) {
    double r[6], dir[6];
    for (int i = 0; i < 6; i++) { x[i] = 0; r[i] = bb[i]; dir[i] = bb[i]; }
    auto dot6 = [](const double* a, const double* b) { double d = 0; for (int i = 0; i < 6; i++) d += a[i] * b[i]; return d; };
    for (int step = 0; step < maxSteps; step++) {
        double Ad[6];
        for (int j = 0; j < 6; j++) { Ad[j] = 0; for (int k = 0; k < 6; k++) Ad[j] += A[j * 6 + k] * dir[k]; }
        double rr = dot6(r, r), Add = dot6(Ad, dir);
        double div = Add < 0 ? -Add : Add;
        double st = 0.0;
        if (div > tol) st = rr / Add;
        for (int i = 0; i < 6; i++) r[i] -= st * Ad[i];
        double rr2 = dot6(r, r);
        for (int i = 0; i < 6; i++) x[i] += st * dir[i];
        if (std::sqrt(rr2) < tol) return;
        double beta = rr2 / rr;
        for (int i = 0; i < 6; i++) dir[i] = r[i] + beta * dir[i];
    }
}
// Textbook Cholesky A = L L^T, forward/back substitution. A non-positive pivot
// (rank-deficient or empty system) yields x = 0, i.e. an identity increment, the
// same outcome as the reference's guarded CG step on a zero system.
inline bool solve_cholesky(const double A[36], const double b[6], double x[6]) {
    double L[36] = {0};
    for (int j = 0; j < 6; j++) {
        double s = A[j * 6 + j];
        for (int k = 0; k < j; k++) s -= L[j * 6 + k] * L[j * 6 + k];
        if (!(s > 0.0)) { for (int i = 0; i < 6; i++) x[i] = 0; return false; }
        double d = std::sqrt(s);
        L[j * 6 + j] = d;
        for (int i = j + 1; i < 6; i++) {
            double t = A[i * 6 + j];
            for (int k = 0; k < j; k++) t -= L[i * 6 + k] * L[j * 6 + k];
            L[i * 6 + j] = t / d;
        }
    }
    double y[6];
    for (int i = 0; i < 6; i++) {
        double t = b[i];
        for (int k = 0; k < i; k++) t -= L[i * 6 + k] * y[k];
        y[i] = t / L[i * 6 + i];
    }
    for (int i = 5; i >= 0; i--) {
        double t = y[i];
        for (int k = i + 1; k < 6; k++) t -= L[k * 6 + i] * x[k];
        x[i] = t / L[i * 6 + i];
    }
    return true;
}

// ---------------------------------------------------------------- A16
// normalizeMat4RealSafe (tracker/kernels/expmCuda.cu:1957-1980): column lengths
// are taken before the cross products and with single-precision sqrtf (Q5).
inline void normalize_mat4_safe(double* T) {
    double u[3] = {T[0], T[4], T[8]}, v[3] = {T[1], T[5], T[9]}, w[3] = {T[2], T[6], T[10]};
    double lenU = (double)sqrtf((float)(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
    double lenW = (double)sqrtf((float)(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]));
    if (lenW > 1e-8f) { w[0] /= lenW; w[1] /= lenW; w[2] /= lenW; }
    else { w[0] = 0; w[1] = 0; w[2] = 1; }
    // crossproductReal(v,w,u)
    u[0] = v[1] * w[2] - v[2] * w[1];
    u[1] = -(v[0] * w[2] - v[2] * w[0]);
    u[2] = v[0] * w[1] - v[1] * w[0];
    if (lenU > 1e-8f) { u[0] /= lenU; u[1] /= lenU; u[2] /= lenU; }
    else { u[0] = 1; u[1] = 0; u[2] = 0; }
    // crossproductReal(w,u,v)
    v[0] = w[1] * u[2] - w[2] * u[1];
    v[1] = -(w[0] * u[2] - w[2] * u[0]);
    v[2] = w[0] * u[1] - w[1] * u[0];
    T[0] = u[0]; T[4] = u[1]; T[8] = u[2];
    T[1] = v[0]; T[5] = v[1]; T[9] = v[2];
    T[2] = w[0]; T[6] = w[1]; T[10] = w[2];
}
// Matrix exponential of the twist matrix A = generateA(x) (warp.cu:910-915) in
// closed form: R = I + a[w]x + b[w]x^2, t = (I + b[w]x + c[w]x^2) v with
// a = sin(th)/th, b = (1-cos th)/th^2, c = (th - sin th)/th^3 (Taylor series for
// small th). The reference evaluates the same exponential with a Pade(6)
// approximant (expmCuda.cu:1989-2048); its outputs are pinned against the
// reference's expm.c in the tests.
inline void se3_exp(const double x[6], double E[16]) {
    double wx = x[0], wy = x[1], wz = x[2];
    double th2 = wx * wx + wy * wy + wz * wz, th = std::sqrt(th2);
    double a, b, c;
    if (th < 1e-4) {
        a = 1.0 - th2 / 6.0 + th2 * th2 / 120.0;
        b = 0.5 - th2 / 24.0 + th2 * th2 / 720.0;
        c = 1.0 / 6.0 - th2 / 120.0 + th2 * th2 / 5040.0;
    } else {
        double s = std::sin(th), co = std::cos(th);
        a = s / th; b = (1.0 - co) / th2; c = (th - s) / (th2 * th);
    }
    double W[9] = {0, -wz, wy, wz, 0, -wx, -wy, wx, 0};
    double W2[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s += W[i * 3 + k] * W[k * 3 + j];
            W2[i * 3 + j] = s;
        }
    double Rm[9], V[9];
    for (int i = 0; i < 9; i++) {
        double I = (i % 4 == 0) ? 1.0 : 0.0;
        Rm[i] = I + a * W[i] + b * W2[i];
        V[i] = I + b * W[i] + c * W2[i];
    }
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) E[i * 4 + j] = Rm[i * 3 + j];
        E[i * 4 + 3] = V[i * 3 + 0] * x[3] + V[i * 3 + 1] * x[4] + V[i * 3 + 2] * x[5];
    }
    E[12] = E[13] = E[14] = 0; E[15] = 1;
}
inline void mat4_mul(const double* A, const double* B, double* C) {
    double t[16];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            t[i * 4 + j] = A[i * 4 + 0] * B[0 * 4 + j] + A[i * 4 + 1] * B[1 * 4 + j] + A[i * 4 + 2] * B[2 * 4 + j] + A[i * 4 + 3] * B[3 * 4 + j];
    std::memcpy(C, t, sizeof(t));
}
// expmKernel (tracker/kernels/expmCuda.cu:1989-2048): Tinc = exp(A(x)),
// normalised, translation times scaleOut; T <- T Tinc, normalised. A zero twist
// leaves Tinc = I.
template <class P>
inline void expm_update(const double x[6], P* TF, double scale) {
    double T[16], Tinc[16];
    for (int i = 0; i < 16; i++) T[i] = TF[i];
    double maxEntry = 0;
    for (int i = 0; i < 6; i++) maxEntry = std::max(maxEntry, std::fabs(x[i]));
    if (maxEntry != 0) {
        se3_exp(x, Tinc);
        normalize_mat4_safe(Tinc);
        Tinc[3] *= scale; Tinc[7] *= scale; Tinc[11] *= scale;
    } else {
        for (int i = 0; i < 16; i++) Tinc[i] = (i % 5 == 0) ? 1.0 : 0.0;
    }
    mat4_mul(T, Tinc, T);
    normalize_mat4_safe(T);
    for (int i = 0; i < 16; i++) TF[i] = (P)T[i];
}

// ---------------------------------------------------------------- A17
// invertRT4 (tracker/tracker/basic_math.cpp:355) and PhotoTracker::applyIncrement
// (tracker/tracker/phototracker.cpp:514-531).
template <class M>
inline void invert_rt4(const M* M4, M* Mi) {
    M t[16];
    t[0] = M4[0]; t[1] = M4[4]; t[2] = M4[8];  t[3] = -(M4[0] * M4[3] + M4[4] * M4[7] + M4[8] * M4[11]);
    t[4] = M4[1]; t[5] = M4[5]; t[6] = M4[9];  t[7] = -(M4[1] * M4[3] + M4[5] * M4[7] + M4[9] * M4[11]);
    t[8] = M4[2]; t[9] = M4[6]; t[10] = M4[10]; t[11] = -(M4[2] * M4[3] + M4[6] * M4[7] + M4[10] * M4[11]);
    t[12] = 0; t[13] = 0; t[14] = 0; t[15] = 1;
    std::memcpy(Mi, t, sizeof(t));
}
// Trel: device pose (ref->cur point map). T: host relative pose (cur->ref),
// Tinc: last increment, global: cumulative pose; all updated in place.
inline void apply_increment(const float* Trel, float* T, double* Tinc, double* global) {
    double Tprev[16], iTprev[16], iT[16], Td[16];
    for (int i = 0; i < 16; i++) Tprev[i] = T[i];
    invert_rt4(Tprev, iTprev);
    for (int i = 0; i < 16; i++) iT[i] = Trel[i];
    invert_rt4(iT, Td);
    for (int i = 0; i < 16; i++) T[i] = (float)Td[i];
    mat4_mul(iTprev, Td, Tinc);
    mat4_mul(global, Tinc, global);
}

// ---------------------------------------------------------------- A24
// PhotoTracker::poseDifferenceExceeded (tracker/tracker/phototracker.cpp:268-290),
// rot2Quaternion (tracker/tracker/basic_math.cpp:943-980, n = 4) and
// normalizeQuaternion (basic_math.cpp:1006-1013).
inline void rot2quat(const float* m, float* q) {
    float tr = m[0] + m[5] + m[10];
    if (tr > 0) {
        float s = 0.5f / sqrtf(tr + 1.0f);
        q[0] = 0.25f / s;
        q[1] = (m[9] - m[6]) * s;
        q[2] = (m[2] - m[8]) * s;
        q[3] = (m[4] - m[1]) * s;
    } else if (m[0] > m[5] && m[0] > m[10]) {
        float s = 2.0f * sqrtf(1.0f + m[0] - m[5] - m[10]);
        q[0] = (m[9] - m[6]) / s; q[1] = 0.25f * s; q[2] = (m[1] + m[4]) / s; q[3] = (m[2] + m[8]) / s;
    } else if (m[5] > m[10]) {
        float s = 2.0f * sqrtf(1.0f + m[5] - m[0] - m[10]);
        q[0] = (m[2] - m[8]) / s; q[1] = (m[1] + m[4]) / s; q[2] = 0.25f * s; q[3] = (m[6] + m[9]) / s;
    } else {
        float s = 2.0f * sqrtf(1.0f + m[10] - m[0] - m[5]);
        q[0] = (m[4] - m[1]) / s; q[1] = (m[2] + m[8]) / s; q[2] = (m[6] + m[9]) / s; q[3] = 0.25f * s;
    }
}
inline bool pose_difference_exceeded(const float* dT, float tthr, float athr) {
    double dx = dT[3], dy = dT[7], dz = dT[11];
    float dist = (float)std::sqrt(dx * dx + dy * dy + dz * dz + 1e-12f);
    if (dist > tthr) return true;
    float angle = 0;
    if ((std::fabs(dT[0] - 1.0f) > 1e-5) || (std::fabs(dT[5] - 1.0f) > 1e-5) || (std::fabs(dT[10] - 1.0f) > 1e-5)) {
        float q[4];
        rot2quat(dT, q);
        float len = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3] + 1e-7f;
        len = (float)std::sqrt((double)len);
        for (int i = 0; i < 4; i++) q[i] /= len;
        double ca = (double)q[0];
        angle = (float)(std::acos(ca) * 2.0f * 180.0f / 3.141592653f);
    }
    return angle > athr;
}

// ---------------------------------------------------------------- A18
// convertZMapToXYZ6Kernel (tracker/kernels/rgb2gray.cu:595-624) and
// computeNormalsCudaKernel (tracker/kernels/d2z.cu:503-537). pts: W*H*6.
template <class R>
void ref_cloud(const R* z, int W, int H, const float* calib, R* pts, int* mask) {
    for (int yi = 0; yi < H; yi++)
        for (int xi = 0; xi < W; xi++) {
            size_t o = (size_t)xi + (size_t)yi * W;
            R p4[4];
            backproject<R>(z[o], xi, yi, calib, p4);
            pts[o * 6 + 0] = p4[0]; pts[o * 6 + 1] = p4[1]; pts[o * 6 + 2] = p4[2];
            mask[o] = p4[3] != 0 ? 1 : 0;
        }
    for (int yi = 0; yi < H; yi++)
        for (int xi = 0; xi < W; xi++) {
            size_t o = (size_t)xi + (size_t)yi * W;
            pts[o * 6 + 3] = 0; pts[o * 6 + 4] = 0; pts[o * 6 + 5] = 0;
            if (xi < 1 || yi < 1 || xi > W - 2 || yi > H - 2) { mask[o] = 0; continue; }
            const R* x1 = &pts[(o - 1) * 6];
            const R* x2 = &pts[(o + 1) * 6];
            const R* y1 = &pts[(o - W) * 6];
            const R* y2 = &pts[(o + W) * 6];
            if (x1[2] < 0 && x2[2] < 0 && y1[2] < 0 && y2[2] < 0) {
                V3<R> u{x2[0] - x1[0], x2[1] - x1[1], x2[2] - x1[2]};
                V3<R> v{y2[0] - y1[0], y2[1] - y1[1], y2[2] - y1[2]};
                V3<R> n;
                n.x = u.y * v.z - u.z * v.y;
                n.y = -(u.x * v.z - u.z * v.x);
                n.z = u.x * v.y - u.y * v.x;
                R len = n.x * n.x + n.y * n.y + n.z * n.z;
                len = (R)1.0f / std::sqrt(len);
                pts[o * 6 + 3] = n.x * len; pts[o * 6 + 4] = n.y * len; pts[o * 6 + 5] = n.z * len;
                continue;
            }
            mask[o] = 0;
        }
}

// genOrientedPoints6Kernel (tracker/kernels/d2z.cu:471-502): the keyframe
// cloud of a filtered reference from ray-cast points and normals (float4
// images); kept where z < 0 and |n|^2 > 1e-7 (FP32 test, as in the kernel),
// every other record zero.
template <typename R>
void oriented_cloud(const float* xyz, const float* nrm, int W, int H, R* pts, int* mask) {
    for (size_t o = 0; o < (size_t)W * H; o++) {
        const float* v = xyz + 4 * o;
        const float* n = nrm + 4 * o;
        const float len = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
        const bool keep = v[2] < 0 && len > 1e-7f;
        for (int k = 0; k < 3; k++) {
            pts[o * 6 + k] = keep ? (R)v[k] : (R)0;
            pts[o * 6 + 3 + k] = keep ? (R)n[k] : (R)0;
        }
        mask[o] = keep ? 1 : 0;
    }
}

// ---------------------------------------------------------------- A19
// genPointsCloud6Kernel (tracker/kernels/d2z.cu:436-469): every z < 0 is valid,
// normal := (0,0,1).
template <class R>
void cur_cloud(const R* z, int W, int H, const float* calib, R* pts, int* mask) {
    R fx = (R)calib[KL_OFFSET], fy = (R)calib[KL_OFFSET + 4];
    R cx = (R)calib[KL_OFFSET + 2], cy = (R)calib[KL_OFFSET + 5];
    for (int yi = 0; yi < H; yi++)
        for (int xi = 0; xi < W; xi++) {
            size_t o = (size_t)xi + (size_t)yi * W;
            R zz = z[o];
            R* p = &pts[o * 6];
            if (zz < 0) {
                p[0] = ((R)xi - cx) * zz / fx;
                p[1] = ((R)yi - cy) * zz / fy;
                p[2] = zz; p[3] = 0; p[4] = 0; p[5] = (R)1.0f;
                mask[o] = 1;
            } else {
                for (int k = 0; k < 6; k++) p[k] = 0;
                mask[o] = 0;
            }
        }
}
// downSampleCloud6Kernel (tracker/kernels/d2z.cu:557-588)
template <class R>
void downsample_cloud(const R* src, const int* msrc, int Wd, int Hd, R* dst, int* mdst) {
    int Ws = Wd * 2;
    for (int yi = 0; yi < Hd; yi++)
        for (int xi = 0; xi < Wd; xi++) {
            size_t o = (size_t)xi + (size_t)yi * Wd;
            size_t h = (size_t)xi * 2 + (size_t)yi * 2 * Ws;
            size_t hs[4] = {h, h + 1, h + Ws, h + Ws + 1};
            R* d = &dst[o * 6];
            for (int k = 0; k < 6; k++) d[k] = 0;
            R n = 0;
            for (int t = 0; t < 4; t++)
                if (msrc[hs[t]]) { n += (R)1; for (int k = 0; k < 6; k++) d[k] += src[hs[t] * 6 + k]; }
            if (n > 0) {
                for (int k = 0; k < 6; k++) d[k] /= n;
                R len = d[3] * d[3] + d[4] * d[4] + d[5] * d[5];
                len = (R)1.0f / std::sqrt(len);
                d[3] *= len; d[4] *= len; d[5] *= len;
                mdst[o] = 1;
            } else mdst[o] = 0;
        }
}

// one channel of bilinearInterpolation6 (tracker/kernels/kernelUtils.h:274-287)
template <class R>
inline R bilerp6(const R* s, R fx, R fy, int pitch) {
    R v0 = s[0], v1 = s[6], v2 = s[pitch], v3 = s[pitch + 6];
    return ((R)1 - fx) * ((R)1 - fy) * v0 + fx * ((R)1 - fy) * v1 + ((R)1 - fx) * fy * v2 + fx * fy * v3;
}

// ---------------------------------------------------------------- A20
// icpResidualKernel2 (tracker/kernels/rgb2gray.cu:223-260): full-image
// point-to-plane residual at level l; all four current taps must be valid.
template <class R, class M>
void icp_residual_mask(const R* ref, const int* rmask, int w, int h, const M* Tm, const float* calib,
                       const R* cur, const int* cmask, int layer, R* res, int* rm) {
    const float* K = calib + KL_OFFSET;
    float af = 1.0f / (float)(1 << layer), bf = 0.5f * (af - 1.0f);
    R a = (R)af, b = (R)bf;
    for (int o = 0; o < w * h; o++) {
        if (rmask[o] == 0) { rm[o] = 0; continue; }
        const R* p = &ref[(size_t)o * 6];
        V3<R> p3{p[0], p[1], p[2]}, n{p[3], p[4], p[5]};
        V3<R> q3 = mul4<R>(Tm, p3);
        V3<R> cn = rot4<R>(Tm, n);
        R qx = q3.x / q3.z, qy = q3.y / q3.z;
        R px = a * ((R)K[0] * qx + (R)K[2]) + b;
        R py = a * ((R)K[4] * qy + (R)K[5]) + b;
        rm[o] = 0;
        if (px >= 0 && px < (R)(w - 1) && py >= 0 && py < (R)(h - 1)) {
            int xdi = (int)px, ydi = (int)py;
            int o2 = xdi + ydi * w;
            if (cmask[o2] && cmask[o2 + 1] && cmask[o2 + w] && cmask[o2 + w + 1]) {
                R fx = px - (R)xdi, fy = py - (R)ydi;
                V3<R> c;
                c.x = bilerp6(&cur[(size_t)o2 * 6 + 0], fx, fy, w * 6);
                c.y = bilerp6(&cur[(size_t)o2 * 6 + 1], fx, fy, w * 6);
                c.z = bilerp6(&cur[(size_t)o2 * 6 + 2], fx, fy, w * 6);
                V3<R> d{c.x - q3.x, c.y - q3.y, c.z - q3.z};
                res[o] = d.x * cn.x + d.y * cn.y + d.z * cn.z;
                rm[o] = 1;
            }
        }
    }
}
// ---------------------------------------------------------------- A21
// selectValidIndex + packIndex (tracker/kernels/rgb2gray.cu:395-468, 827-850):
// intended semantics (Q6) = ascending-index compaction of mask > 0.
inline int compact(const int* mask, int n, int* out) {
    int k = 0;
    for (int i = 0; i < n; i++)
        if (mask[i] > 0) out[k++] = i;
    return k;
}

// ---------------------------------------------------------------- A22
// generateStudentTWeights (tracker/kernels/rgb2gray.cu:470-568, 854-877):
// w = 6/(5 + r^2/sigma^2), sigma0^2 = 100; three rounds of S = tree-sum of
// (int)(w_i r_i^2) over 1024-element blocks (one element per thread, so the int
// accumulator truncates each product, Q7), blocks combined by the last block
// with the same tree; S = max(S, N) only when more than one block ran; then
// sigma^2 = S/N; finally w = min(w, 1).
template <class R>
inline R tree1024(R* p) {
    for (int s = 512; s >= 1; s >>= 1)
        for (int t = 0; t < s; t++) p[t] += p[t + s];
    return p[0];
}
template <class R>
inline void student_t_weights(const R* r2, int n, R* w, R* sigma2_out = nullptr) {
    for (int i = 0; i < n; i++) w[i] = (R)6 / ((R)5 + r2[i] / (R)100.0f);
    int nb = (n + 1023) / 1024;
    R var = 0;
    for (int it = 0; it < 3; it++) {
        std::vector<R> blockSum(nb);
        std::vector<R> p(1024);
        for (int bI = 0; bI < nb; bI++) {
            for (int t = 0; t < 1024; t++) {
                int i = bI * 1024 + t;
                int sum = 0;
                if (i < n) sum = (int)((R)sum + r2[i] * w[i]);
                p[t] = (R)sum;
            }
            blockSum[bI] = tree1024(p.data());
        }
        if (nb == 1) var = blockSum[0];
        else {
            for (int t = 0; t < 1024; t++) {
                int sum = 0;
                for (int i = t; i < nb; i += 1024) sum = (int)((R)sum + blockSum[i]);
                p[t] = (R)sum;
            }
            var = tree1024(p.data());
            R minVal = (R)n;
            if (var < minVal) var = minVal;
        }
        for (int i = 0; i < n; i++) w[i] = (R)6 / ((R)5 + r2[i] / (var / (R)n));
    }
    for (int i = 0; i < n; i++) w[i] = std::min(w[i], (R)1.0f);
    if (sigma2_out) *sigma2_out = var / (R)n;
}

// ---------------------------------------------------------------- A23
// icpResidualIndexKernel2 (tracker/kernels/rgb2gray.cu:323-390): weighted, scaled
// point-to-plane residual and its 6 Jacobian columns for the compacted points.
template <class R, class M>
void icp_residual_jacobian(const R* ref, const int* sel, int n, int w, int h, int layer, const M* Tm,
                           const float* calib, const R* cur, const int* cmask, const R* wts,
                           float optScaleIn, R* res, R* const* J) {
    const float* K = calib + KL_OFFSET;
    float af = 1.0f / (float)(1 << layer), bf = 0.5f * (af - 1.0f);
    R a = (R)af, b = (R)bf, s = (R)optScaleIn;
    for (int i = 0; i < n; i++) {
        const R* p = &ref[(size_t)sel[i] * 6];
        V3<R> p3{p[0], p[1], p[2]}, rn{p[3], p[4], p[5]};
        V3<R> q3 = mul4<R>(Tm, p3);
        V3<R> cn = rot4<R>(Tm, rn);
        R qx = q3.x / q3.z, qy = q3.y / q3.z;
        R px = a * ((R)K[0] * qx + (R)K[2]) + b;
        R py = a * ((R)K[4] * qy + (R)K[5]) + b;
        res[i] = 0;
        for (int k = 0; k < 6; k++) J[k][i] = 0;
        if (px >= 0 && px < (R)(w - 1) && py >= 0 && py < (R)(h - 1)) {
            R wt = (R)wts[i];
            int xdi = (int)px, ydi = (int)py;
            R fx = px - (R)xdi, fy = py - (R)ydi;
            int o2 = xdi + ydi * w;
            V3<R> c;
            c.x = bilerp6(&cur[(size_t)o2 * 6 + 0], fx, fy, w * 6);
            c.y = bilerp6(&cur[(size_t)o2 * 6 + 1], fx, fy, w * 6);
            c.z = bilerp6(&cur[(size_t)o2 * 6 + 2], fx, fy, w * 6);
            if (cmask[o2] && wt > 0 && std::fabs(c.z - q3.z) < (R)300.0f) {
                V3<R> d{c.x - q3.x, c.y - q3.y, c.z - q3.z};
                res[i] = s * wt * (d.x * cn.x + d.y * cn.y + d.z * cn.z);
                V3<R> gp[3] = {{0, -p3.z, p3.y}, {p3.z, 0, -p3.x}, {-p3.y, p3.x, 0}};
                V3<R> gn[3] = {{0, -rn.z, rn.y}, {rn.z, 0, -rn.x}, {-rn.y, rn.x, 0}};
                for (int k = 0; k < 3; k++) {
                    V3<R> dB = rot4<R>(Tm, gp[k]);
                    V3<R> dN = rot4<R>(Tm, gn[k]);
                    J[k][i] = -s * wt * (-cn.x * dB.x - cn.y * dB.y - cn.z * dB.z + d.x * dN.x + d.y * dN.y + d.z * dN.z);
                }
                V3<R> e[3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
                for (int k = 0; k < 3; k++) {
                    V3<R> dB = rot4<R>(Tm, e[k]);
                    J[3 + k][i] = -wt * (-cn.x * dB.x - cn.y * dB.y - cn.z * dB.z);
                }
            }
        }
    }
}

}  // namespace oracle
