// Device-side geometry and sampling helpers of the tracker kernels.
//
// Expressions keep the operand order of the reference's device helpers
// (tracker/kernels/kernelUtils.h) and the library is compiled with
// -ffp-contract=off, so every FP32 value is the correctly rounded result of the
// same sequence of operations as in the reference source.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rt {

// calibration block offsets (tracker/calib/calib.h:20-36)
enum {
    KR = 0, IKR = 9, KCR = 18, TLR = 26, C1 = 42, KL = 43, IKL = 52, C0 = 61, MIND = 62, MAXD = 63, TRL = 64,
    CALIB = 96
};

struct f3 { float x, y, z; };
struct f2 { float x, y; };

// matrixMultVec4 (kernelUtils.h:242-247)
__device__ __forceinline__ f3 mul4(const float* T, f3 v) {
    f3 r;
    r.x = T[0] * v.x + T[1] * v.y + T[2] * v.z + T[3];
    r.y = T[4] * v.x + T[5] * v.y + T[6] * v.z + T[7];
    r.z = T[8] * v.x + T[9] * v.y + T[10] * v.z + T[11];
    return r;
}
// matrixRot4 (kernelUtils.h:249-254)
__device__ __forceinline__ f3 rot4(const float* T, f3 v) {
    f3 r;
    r.x = T[0] * v.x + T[1] * v.y + T[2] * v.z;
    r.y = T[4] * v.x + T[5] * v.y + T[6] * v.z;
    r.z = T[8] * v.x + T[9] * v.y + T[10] * v.z;
    return r;
}
// matrixMultVec3 (kernelUtils.h:235-240)
__device__ __forceinline__ f3 mul3(const float* K, f3 v) {
    f3 r;
    r.x = K[0] * v.x + K[1] * v.y + K[2] * v.z;
    r.y = K[3] * v.x + K[4] * v.y + K[5] * v.z;
    r.z = K[6] * v.x + K[7] * v.y + K[8] * v.z;
    return r;
}
// distortPoint (kernelUtils.h:394-408)
__device__ __forceinline__ f2 distort(f2 pu, const float* kc, const float* K) {
    float dx = pu.x * pu.x, dy = pu.y * pu.y;
    float r2 = dx + dy, r4 = r2 * r2, r6 = r4 * r2;
    float rd = 1.0f + kc[0] * r2 + kc[1] * r4 + kc[4] * r6;
    f2 pd;
    pd.x = K[0] * pu.x * rd + K[2];
    pd.y = K[4] * pu.y * rd + K[5];
    return pd;
}
// inBounds (kernelUtils.h:341-344)
__device__ __forceinline__ bool in_bounds(f2 p, float xmax, float ymax) {
    return p.x >= 10.0f && p.x < xmax && p.y >= 10.0f && p.y < ymax;
}
__device__ __forceinline__ bool in_bounds(f2 p, int w, int h) { return in_bounds(p, (float)(w - 10), (float)(h - 10)); }
// A calibration whose colour camera coincides with the depth camera: TLR and
// TRL are identities, no radial distortion (k1 = k2 = k5 = 0), KR = KL and a
// pinhole without skew. Every transform skipped under it is an exact identity
// (x*1 + y*0 + z*0 + 0 rounds to x, 1 + 0*r2 + 0*r4 + 0*r6 to 1, and the
// depth-camera reprojection of a colour-camera point is that point), so the
// kernels may take the shorter path and compute the same values; the only
// possible difference, the sign of a zero coordinate, is mapped to the same
// pixel position by every later use (cx, cy are nonzero).
__host__ __device__ __forceinline__ bool calib_simple(const float* c) {
    bool s = c[KCR] == 0.0f && c[KCR + 1] == 0.0f && c[KCR + 4] == 0.0f;
#pragma unroll
    for (int k = 0; k < 12; k++) {
        const float id = (k % 5 == 0) ? 1.0f : 0.0f;
        s = s && c[TLR + k] == id && c[TRL + k] == id;
    }
#pragma unroll
    for (int k = 0; k < 9; k++) s = s && c[KR + k] == c[KL + k];
    return s && c[KR + 1] == 0.0f && c[KR + 3] == 0.0f && c[KR + 6] == 0.0f && c[KR + 7] == 0.0f &&
           c[KR + 8] == 1.0f && c[KR + 2] != 0.0f && c[KR + 5] != 0.0f;
}
// distortPoint with zero distortion: K applied to the normalised point
__device__ __forceinline__ f2 project_simple(f2 pu, const float* K) {
    return f2{K[0] * pu.x + K[2], K[4] * pu.y + K[5]};
}

// Pointers into device memory (global address space): loads through them are
// global_load with a scalar base, not flat loads, which would also wait on the
// LDS counter and need 64-bit addresses per lane.
#define RT_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ const RT_GLOBAL T* gmem(const T* p) { return (const RT_GLOBAL T*)p; }
template <class T>
__device__ __forceinline__ RT_GLOBAL T* gmem(T* p) { return (RT_GLOBAL T*)p; }

typedef float nfloat4 __attribute__((ext_vector_type(4)));
typedef float nfloat2 __attribute__((ext_vector_type(2)));
// element i of a device array, loaded through the global address space
template <class T>
__device__ __forceinline__ T ldg(const T* p, size_t i) { return gmem(p)[i]; }
// the element at byte offset off (32 bits) of a device array: one load with a
// scalar base address and a 32-bit offset per lane (no 64-bit address arithmetic)
template <class T>
__device__ __forceinline__ T at_offset(const RT_GLOBAL T* p, unsigned off) {
    return *(const RT_GLOBAL T*)((const RT_GLOBAL char*)p + off);
}
template <class T>
__device__ __forceinline__ void store_at_offset(RT_GLOBAL T* p, unsigned off, T v) {
    *(RT_GLOBAL T*)((RT_GLOBAL char*)p + off) = v;
}
__device__ __forceinline__ float4 ldg(const float4* p, size_t i) {
    nfloat4 v = ((const RT_GLOBAL nfloat4*)p)[i];
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float2 ldg(const float2* p, size_t i) {
    nfloat2 v = ((const RT_GLOBAL nfloat2*)p)[i];
    return make_float2(v.x, v.y);
}

// stores of element i through the global address space
__device__ __forceinline__ void stg(float4* p, size_t i, float4 v) {
    ((RT_GLOBAL nfloat4*)p)[i] = nfloat4{v.x, v.y, v.z, v.w};
}
__device__ __forceinline__ void stg(float2* p, size_t i, float2 v) { ((RT_GLOBAL nfloat2*)p)[i] = nfloat2{v.x, v.y}; }
__device__ __forceinline__ void stg(float* p, size_t i, float v) { gmem(p)[i] = v; }

// Image sources: FP32 planes, or the raw inputs read through their exact
// conversions: 8-bit gray as (float)g/255 from a 256-entry table, 16-bit depth
// in mm as -(float)d (the values the conversion kernel would have stored).
struct F32Src {
    const RT_GLOBAL float* __restrict__ p;
    __device__ F32Src(const float* q) : p(gmem(q)) {}
    __device__ __forceinline__ float operator[](size_t i) const { return p[i]; }
};
struct U8Src {
    const RT_GLOBAL uint8_t* __restrict__ p;
    const float* lut;
    __device__ U8Src(const uint8_t* q, const float* l) : p(gmem(q)), lut(l) {}
    __device__ __forceinline__ float operator[](size_t i) const { return lut[p[i]]; }
};
struct U16DepthSrc {
    const RT_GLOBAL uint16_t* __restrict__ p;
    __device__ U16DepthSrc(const uint16_t* q) : p(gmem(q)) {}
    __device__ __forceinline__ float operator[](size_t i) const { return -(float)p[i]; }
};
// A 16-bit depth map stored in 8 x 8-pixel blocks of 64 values, block
// (x / 8, y / 8) at (y / 8) (W / 8) + x / 8, row-major inside: one 128-byte
// line per block, so that nearby taps share lines in both directions. The same
// negated values as U16DepthSrc.
struct U16BlockDepthSrc {
    const RT_GLOBAL uint16_t* __restrict__ p;
};
// the depth sample at pixel (x, y) of a W-wide depth source
template <class ZS>
__device__ __forceinline__ float depth_tap(const ZS& z, int x, int y, int W) { return z[x + (size_t)y * W]; }
__device__ __forceinline__ float depth_tap(const U16BlockDepthSrc& z, int x, int y, int W) {
    const unsigned blk = __umul24((unsigned)y >> 3, (unsigned)W >> 3) + ((unsigned)x >> 3);
    return -(float)z.p[(blk << 6) + ((y & 7) << 3) + (x & 7)];
}

// bilinearInterpolation(xdi,ydi,fx,fy,...) (kernelUtils.h:267-272)
__device__ __forceinline__ float blend(float v0, float v1, float v2, float v3, float fx, float fy) {
    return (1.0f - fx) * (1.0f - fy) * v0 + fx * (1.0f - fy) * v1 + (1.0f - fx) * fy * v2 + fx * fy * v3;
}
template <class S>
__device__ __forceinline__ float bilerp(const S& img, int width, int xdi, int ydi, float fx, float fy) {
    size_t o = (size_t)xdi + (size_t)ydi * width;
    return blend(img[o], img[o + 1], img[o + width], img[o + width + 1], fx, fy);
}
__device__ __forceinline__ float bilerp(const float* __restrict__ img, int width, int xdi, int ydi, float fx, float fy) {
    return bilerp(F32Src{img}, width, xdi, ydi, fx, fy);
}
// convertZMapToXYZKernel (rgb2gray.cu:570-593): back-projection with KL,
// valid iff -max < z < -min. Returns validity; p = 0 when invalid.
__device__ __forceinline__ bool backproject(float z, int xi, int yi, const float* c, f3& p) {
    float mn = c[MIND], mx = c[MAXD];
    if (z < -mn && z > -mx) {
        float fx = c[KL], fy = c[KL + 4], cx = c[KL + 2], cy = c[KL + 5];
        p.x = ((float)xi - cx) * z / fx;
        p.y = ((float)yi - cy) * z / fy;
        p.z = z;
        return true;
    }
    p.x = 0.0f; p.y = 0.0f; p.z = 0.0f;
    return false;
}

// Level-0 photometric features of one keyframe pixel (xyz2CloudKernel,
// d2z.cu:235-331): point, gray and central-difference gradient. A pixel without
// a valid projection gets the reference's empty record: point (0,0,-maxDist),
// zero gray and gradient.
struct PixFeat { f3 p; float c1, gx, gy, gmag; bool ok; };
template <class ZS, class GS>
__device__ __forceinline__ PixFeat pixel_features(const ZS& z, const GS& g0, int W, int H, int xi, int yi,
                                                  const float* c) {
    PixFeat f;
    f3 p3;
    backproject(z[(size_t)xi + (size_t)yi * W], xi, yi, c, p3);
    if (p3.z < 0.0f) {
        f3 r3 = mul4(c + TLR, p3);
        f2 pu{r3.x / r3.z, r3.y / r3.z};
        f2 p = distort(pu, c + KCR, c + KR);
        if (in_bounds(p, W, H)) {
            int xdi = (int)p.x, ydi = (int)p.y;
            float fx = p.x - (float)xdi, fy = p.y - (float)ydi;
            // The five samples read a 4x4 neighbourhood; load its 12 used taps once.
            // The reference indexes the shifted samples as (int)(p +- 1.0f),
            // which equals xdi +- 1 unless p sits just below a power of two
            // where the sum rounds up: those samples are read separately.
            int xe = (int)(p.x + 1.0f), xw = (int)(p.x - 1.0f), yn = (int)(p.y - 1.0f), ys = (int)(p.y + 1.0f);
            size_t o = (size_t)xdi + (size_t)ydi * W;
            float a01 = g0[o - W], a02 = g0[o - W + 1];
            float a10 = g0[o - 1], a11 = g0[o], a12 = g0[o + 1], a13 = g0[o + 2];
            float a20 = g0[o + W - 1], a21 = g0[o + W], a22 = g0[o + W + 1], a23 = g0[o + W + 2];
            float a31 = g0[o + 2 * W], a32 = g0[o + 2 * W + 1];
            // the four shifted blends are formed before the rare re-reads, so
            // that all twelve loads are issued together
            float c1 = blend(a11, a12, a21, a22, fx, fy);
            float bN = blend(a01, a02, a11, a12, fx, fy), bE = blend(a12, a13, a22, a23, fx, fy);
            float bW = blend(a10, a11, a20, a21, fx, fy), bS = blend(a21, a22, a31, a32, fx, fy);
            float cN = yn == ydi - 1 ? bN : bilerp(g0, W, xdi, yn, fx, fy);
            float cE = xe == xdi + 1 ? bE : bilerp(g0, W, xe, ydi, fx, fy);
            float cW = xw == xdi - 1 ? bW : bilerp(g0, W, xw, ydi, fx, fy);
            float cS = ys == ydi + 1 ? bS : bilerp(g0, W, xdi, ys, fx, fy);
            f.gx = (cE - cW) / 2.0f;
            f.gy = (cS - cN) / 2.0f;
            f.gmag = fminf(fabsf(f.gy) + fabsf(f.gx), 1.0f);
            f.c1 = c1;
            f.p = p3;
            f.ok = true;
            return f;
        }
    }
    f.p = f3{0.0f, 0.0f, -c[MAXD]};
    f.c1 = 0.0f; f.gx = 0.0f; f.gy = 0.0f; f.gmag = 0.0f; f.ok = false;
    return f;
}
__device__ __forceinline__ PixFeat pixel_features(const float* __restrict__ z, const float* __restrict__ g0, int W,
                                                  int H, int xi, int yi, const float* c) {
    return pixel_features(F32Src{z}, F32Src{g0}, W, H, xi, yi, c);
}

// a / b rounded to nearest, given y = RN(1/b) (an IEEE division, done once per
// divisor): q = RN(a y), then one fused correction with the exact remainder
// r = a - b q, q' = RN(q + r y). By Markstein's theorem (P. Markstein, IBM J.
// Res. Dev. 34(1), 1990; Muller et al., Handbook of Floating-Point
// Arithmetic, division with an FMA) q' is the correctly rounded quotient
// when no intermediate overflows or underflows, so it equals the IEEE
// division bit for bit, except that a = -0 over a positive b gives +0 (the
// division gives -0). Three instructions instead of about eleven.
__device__ __forceinline__ float div_rn(float a, float b, float y) {
    const float q = a * y;
    const float r = __builtin_fmaf(-b, q, a);
    return __builtin_fmaf(r, y, q);
}
// an 8-bit gray value t as (float)t / 255, the IEEE quotient, by div_rn
// (three instructions; all 256 values checked against IEEE division)
__device__ __forceinline__ float u8_unit(unsigned t) { return div_rn((float)t, 255.0f, 1.0f / 255.0f); }
// RN(1 / z) for z = -u, u an integer in [1, 65535] (a negated 16-bit depth
// value): the hardware reciprocal (within 1 ulp) refined by one Newton step
// with fused multiply-adds, y + y (1 - z y). The residual 1 - z y is exact,
// and for these 65535 inputs the result equals the IEEE division 1.0f / z bit
// for bit (compared exhaustively against IEEE float32 division by a test); it
// is not claimed for other inputs. Three instructions instead of about eleven.
__device__ __forceinline__ float rcp_depth(float z) {
    const float y = __builtin_amdgcn_rcpf(z);
    const float e = __builtin_fmaf(-z, y, 1.0f);
    return __builtin_fmaf(e, y, y);
}
// Per-image constants of the gradient pass: RN(1/fx), RN(1/fy) of the depth
// camera (KL) for div_rn, and the in-bounds limits (float)(W - 10),
// (float)(H - 10)
struct GradConst { float ifx, ify, xmax, ymax; };

// A block of gray values staged in LDS: image pixels [x0, x0+w) x [y0, y0+h),
// row pitch w, holding exactly the values the image source returns.
struct LdsTile {
    const float* p;
    int x0, y0, w, h;
    __device__ __forceinline__ bool covers(int xa, int ya, int xb, int yb) const {
        return xa >= x0 && ya >= y0 && xb < x0 + w && yb < y0 + h;
    }
};
// The gradient magnitude of pixel_features, with the 4x4 tap neighbourhood
// read from an LDS tile when it lies inside it (the same values, so the same
// results), else from the image; zc is the pixel's own depth value
// z[xi + yi*W], loaded by the caller: the negated 16-bit depth, an integer.
// The four divisions of the back-projection and the projection run as div_rn
// (the same quotients; a zero coordinate may change sign, which the
// projection maps to the same position: it is only ever added to cx, cy or
// scaled into a sum). SIMPLE: the calibration satisfies calib_simple; then
// the projection divides by zc itself, whose reciprocal rcp_depth gives.
// In bounds, p.x >= 10, so p.x - 1 is exact and (int)(p.x - 1) = xdi - 1
// (likewise for y): of the four neighbour positions only p + 1 can round
// onto the next integer, and only xe and ys are tested.
template <bool SIMPLE, class ZS, class GS>
__device__ __forceinline__ float grad_mag_tiled(float zc, const ZS& z, const GS& g0, const LdsTile& tile, int W,
                                                int H, int xi, int yi, const float* c, GradConst k) {
    f3 p3{0.0f, 0.0f, 0.0f};
    if (zc < -c[MIND] && zc > -c[MAXD]) {  // backproject
        p3.x = div_rn(((float)xi - c[KL + 2]) * zc, c[KL], k.ifx);
        p3.y = div_rn(((float)yi - c[KL + 5]) * zc, c[KL + 4], k.ify);
        p3.z = zc;
    }
    if (p3.z < 0.0f) {
        f3 r3 = SIMPLE ? p3 : mul4(c + TLR, p3);
        const float iz = SIMPLE ? rcp_depth(r3.z) : 1.0f / r3.z;
        f2 pu{div_rn(r3.x, r3.z, iz), div_rn(r3.y, r3.z, iz)};
        f2 p = SIMPLE ? project_simple(pu, c + KR) : distort(pu, c + KCR, c + KR);
        if (in_bounds(p, k.xmax, k.ymax)) {
            int xdi = (int)p.x, ydi = (int)p.y;
            float fx = p.x - (float)xdi, fy = p.y - (float)ydi;
            int xe = (int)(p.x + 1.0f), ys = (int)(p.y + 1.0f);
            if (xe == xdi + 1 && ys == ydi + 1 && tile.covers(xdi - 1, ydi - 1, xdi + 2, ydi + 2)) {
                const float* q = tile.p + __umul24(ydi - tile.y0, tile.w) + (xdi - tile.x0);  // both >= 1: covered
                const int P = tile.w;
                float a01 = q[-P], a02 = q[-P + 1];
                float a10 = q[-1], a11 = q[0], a12 = q[1], a13 = q[2];
                float a20 = q[P - 1], a21 = q[P], a22 = q[P + 1], a23 = q[P + 2];
                float a31 = q[2 * P], a32 = q[2 * P + 1];
                float cN = blend(a01, a02, a11, a12, fx, fy);
                float cE = blend(a12, a13, a22, a23, fx, fy);
                float cW = blend(a10, a11, a20, a21, fx, fy);
                float cS = blend(a21, a22, a31, a32, fx, fy);
                float gx = (cE - cW) / 2.0f;
                float gy = (cS - cN) / 2.0f;
                return fminf(fabsf(gy) + fabsf(gx), 1.0f);
            }
            return pixel_features(z, g0, W, H, xi, yi, c).gmag;
        }
    }
    return 0.0f;
}

// gradient-magnitude bin (gpu_hist.cu:135, filterIndex256.cu:126), fused multiply-add
__device__ __forceinline__ unsigned grad_bin(float g) { return (unsigned)__builtin_fmaf(g, 255.0f, 0.5f); }

// wave-uniform copy of a value (all lanes hold the same): lives in a scalar register
__device__ __forceinline__ float uni(float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); }
// The 96-float calibration block, staged in LDS by the workgroup, as
// wave-uniform values: the helpers index it with constants, so it stays in
// scalar registers instead of being loaded per pixel.
__device__ __forceinline__ void uniform_calib(const float* cal_lds, float* cl) {
#pragma unroll
    for (int k = 0; k < 96; k++) cl[k] = uni(cal_lds[k]);
}

// wave64 helpers
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }
// Sums of 32 per-lane values over the 64 lanes of a wave, by halving: at the
// step with lane offset o every lane keeps half of its remaining values (the
// upper half when lane & o) and adds its partner's copy of that half, so 32
// values need 16+8+4+2+1+1 exchanges instead of 32 six-step trees. The first
// two steps are v_permlane32_swap / v_permlane16_swap, which exchange lanes
// 32-63 of one register with lanes 0-31 of another (rows 1, 3 with rows 0, 2
// of 16 lanes): after the swap the two registers hold, lane by lane, the
// kept value and the partner's copy of it. On return lanes 2k and 2k+1 hold
// the total of value k.
__device__ __forceinline__ float swap_add32(float x, float y) {
    auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(y), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ __forceinline__ float swap_add16(float x, float y) {
    auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(y), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ __forceinline__ float wave_sum32(const float* v) {
    const int lane = lane_id();
    float a[16];
#pragma unroll
    for (int j = 0; j < 16; j++) a[j] = swap_add32(v[j], v[16 + j]);   // lane & 32 ? value 16 + j : value j
#pragma unroll
    for (int j = 0; j < 8; j++) a[j] = swap_add16(a[j], a[8 + j]);     // lane & 16 ? 8 + j : j
#pragma unroll
    for (int n = 4; n >= 1; n >>= 1) {  // offsets 8, 4, 2: both sums, then a select of values
        const bool up = lane & (2 * n);
#pragma unroll
        for (int j = 0; j < n; j++) {
            const float x = a[j], y = a[n + j];
            const float sx = x + __shfl_xor(x, 2 * n, 64), sy = y + __shfl_xor(y, 2 * n, 64);
            a[j] = up ? sy : sx;
        }
    }
    return a[0] + __shfl_xor(a[0], 1, 64);
}
// The same for 32 FP64 values: each exchange moves the two halves of a double
// through the same lane swap.
__device__ __forceinline__ double swap_add32d(double x, double y) {
    const unsigned long long xb = __double_as_longlong(x), yb = __double_as_longlong(y);
    auto lo = __builtin_amdgcn_permlane32_swap((unsigned)xb, (unsigned)yb, false, false);
    auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(xb >> 32), (unsigned)(yb >> 32), false, false);
    const double a = __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0]));
    const double b = __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
    return a + b;
}
__device__ __forceinline__ double swap_add16d(double x, double y) {
    const unsigned long long xb = __double_as_longlong(x), yb = __double_as_longlong(y);
    auto lo = __builtin_amdgcn_permlane16_swap((unsigned)xb, (unsigned)yb, false, false);
    auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(xb >> 32), (unsigned)(yb >> 32), false, false);
    const double a = __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0]));
    const double b = __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
    return a + b;
}
__device__ __forceinline__ double wave_sum32d(const double* v) {
    const int lane = lane_id();
    double a[16];
#pragma unroll
    for (int j = 0; j < 16; j++) a[j] = swap_add32d(v[j], v[16 + j]);
#pragma unroll
    for (int j = 0; j < 8; j++) a[j] = swap_add16d(a[j], a[8 + j]);
#pragma unroll
    for (int n = 4; n >= 1; n >>= 1) {
        const bool up = lane & (2 * n);
#pragma unroll
        for (int j = 0; j < n; j++) {
            const double x = a[j], y = a[n + j];
            const double sx = x + __shfl_xor(x, 2 * n, 64), sy = y + __shfl_xor(y, 2 * n, 64);
            a[j] = up ? sy : sx;
        }
    }
    return a[0] + __shfl_xor(a[0], 1, 64);
}

}  // namespace rt
