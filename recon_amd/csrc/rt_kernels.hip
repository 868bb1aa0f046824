// HIP kernels of the single-pair tracker path (gfx950, wave64).
//
// Keyframe preparation: one fused convert+pyramid pass, a per-pixel feature
// pass that also builds the 256-bin gradient histogram in LDS, a three-kernel
// ordered selection (segment counts, one-workgroup scan with the reference's
// trim rule, ordered scatter), and one gather pass that writes each selected
// point's position and its per-level Jacobian records.
// Gauss-Newton iteration: a residual pass (bilinear warp, zWeight, MAD
// histogram in LDS), a system pass (robust weights and the 27 unique
// normal-equation terms: FP32 products summed in FP64 per workgroup) and a
// one-workgroup pass that combines the workgroups' sums in FP64 and runs the
// FP64 solve and the SE(3) update. The geometric side adds an ordered
// compaction of the valid ICP pixels, three Student-t variance rounds and its
// own system pass, which a bi-objective iteration runs beside the residuals.
#include <hip/hip_runtime.h>
#include <math.h>
#include "rt_device.hpp"
#include "rt_kernels.hpp"
#include "rt_prep.hpp"

namespace rt {

// ------------------------------------------------------------------ A1 + A2
// rgbTex2GrayHdrZKernel (rgb2gray.cu:64-76) / u8-u16 conversion, fused with
// fastDownSampleHdrKernel (downSample2Cuda.cu:35-42) for all levels. Each
// thread owns a 4x4 level-0 block (its 2x2 level-1 block and one level-2
// pixel); level 3 comes from the workgroup's 16x16 level-2 tile in LDS.
// src: 0 = u8 gray + u16 depth, 1 = float4 RGBD (bottom-up), 2 = level 0
// already in place (rebuild levels >= 1 only).
template <int SRC>
__global__ __launch_bounds__(256) void k_convert_pyramid(const uint8_t* __restrict__ g8,
                                                         const uint16_t* __restrict__ d16,
                                                         const float4* __restrict__ rgbd, int W, int H, int L,
                                                         const float* __restrict__ c, float* __restrict__ g0,
                                                         float* __restrict__ g1, float* __restrict__ g2,
                                                         float* __restrict__ g3, float* __restrict__ z,
                                                         unsigned* zero_a, int na, unsigned* zero_b, int nb) {
    convert_pyramid_block<SRC>(blockIdx.x, blockIdx.y, g8, d16, rgbd, W, H, L, c, g0, g1, g2, g3, z);
    if (blockIdx.x == 0 && blockIdx.y == 0) {
        for (int i = threadIdx.x; i < na; i += 256) zero_a[i] = 0u;
        for (int i = threadIdx.x; i < nb; i += 256) zero_b[i] = 0u;
    }
}

void launch_convert_pyramid(const Dev& d, const Pyr& p, const uint8_t* g8, const uint16_t* d16, const float* rgbd,
                            bool zero_madhist, bool zero_ghist, hipStream_t s) {
    dim3 grid((d.W / 4 + 15) / 16, (d.H / 4 + 15) / 16);
    unsigned* za = zero_madhist ? d.madhist : nullptr;
    unsigned* zb = zero_ghist ? d.ghist : nullptr;
    int na = zero_madhist ? 256 : 0, nb = zero_ghist ? 256 : 0;
    if (g8)
        k_convert_pyramid<0><<<grid, 256, 0, s>>>(g8, d16, nullptr, d.W, d.H, d.L, d.calib, p.gray[0], p.gray[1],
                                                  p.gray[2], p.gray[3], p.z, za, na, zb, nb);
    else if (rgbd)
        k_convert_pyramid<1><<<grid, 256, 0, s>>>(nullptr, nullptr, reinterpret_cast<const float4*>(rgbd), d.W, d.H,
                                                  d.L, d.calib, p.gray[0], p.gray[1], p.gray[2], p.gray[3], p.z, za, na,
                                                  zb, nb);
    else
        k_convert_pyramid<2><<<grid, 256, 0, s>>>(nullptr, nullptr, nullptr, d.W, d.H, d.L, d.calib, p.gray[0],
                                                  p.gray[1], p.gray[2], p.gray[3], p.z, za, na, zb, nb);
}

// Clears the relative pose: TrelDev and, for sequences, the device copy of the
// host relative pose (PhotoTracker::prepareReference, phototracker.cpp:363-369).
__global__ void k_reset_pose(float* T, double* pose) {
    int i = threadIdx.x;
    if (i < 16) {
        T[i] = (i % 5 == 0) ? 1.0f : 0.0f;
        if (pose) pose[i] = (i % 5 == 0) ? 1.0 : 0.0;
    }
}
void launch_reset_pose(float* T, double* pose, hipStream_t s) { k_reset_pose<<<1, 64, 0, s>>>(T, pose); }

__global__ void k_fill(float* p, float v, int n) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}
void launch_fill(float* p, float v, int n, hipStream_t s) { k_fill<<<(n + 255) / 256, 256, 0, s>>>(p, v, n); }

// ------------------------------------------------------------------ A3-A6
// xyz2CloudKernel (d2z.cu:235-331) reduced to the gradient magnitude the
// selection needs, plus the 256-bin histogram (cuda_hist.cu:100-152) built in
// LDS and added to the global histogram once per workgroup.
__global__ __launch_bounds__(256) void k_ref_features(const float* __restrict__ z, const float* __restrict__ g0, int W,
                                                      int H, const float* __restrict__ c, float* __restrict__ gmag,
                                                      unsigned* __restrict__ ghist) {
    ref_features_block(blockIdx.x, blockIdx.y, z, g0, W, H, c, gmag, ghist);
}

// ------------------------------------------------------------------ A7
__global__ __launch_bounds__(256) void k_select_count(const float* __restrict__ gmag, int npix, const unsigned* ghist,
                                                      int N, int* seg) {
    select_count_block(blockIdx.x, gmag, npix, ghist, N, seg);
}
__global__ __launch_bounds__(512) void k_select_scan(int* seg) { select_scan_block(seg); }
__global__ __launch_bounds__(256) void k_select_scatter(const float* __restrict__ gmag, int npix, int N,
                                                        const int* __restrict__ seg, int* __restrict__ sel) {
    select_scatter_block(blockIdx.x, gmag, npix, N, seg, sel);
}

// ------------------------------------------------------------------ A8 + A9
// addVertexAttributesWithoutNormalsKernel3/4 (d2z.cu:700-936) and
// precomputeJacobian4Kernel (warp.cu:502-645) for the selected points. Per
// point and level the record holds {J0,J1,J2,J3} and {J4,J5,I_ref,0}.
__global__ __launch_bounds__(256) void k_ref_points(const float* __restrict__ z, const float* __restrict__ g0,
                                                    const float* __restrict__ g1, const float* __restrict__ g2,
                                                    const float* __restrict__ g3, int W, int H, int L, int N,
                                                    const float* __restrict__ c, float scale_in,
                                                    const int* __restrict__ sel, float4* __restrict__ pts,
                                                    float4* __restrict__ j0, float4* __restrict__ j1,
                                                    float4* __restrict__ j2, float4* __restrict__ j3) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    ref_point(i, F32Src{z}, F32Src{g0}, g1, g2, g3, W, H, L, N, c, scale_in, sel, pts, j0, j1, j2, j3);
}

// The keyframe's depth for display: xyz2DepthNormalizedKernel (rgb2gray.cu:104-116)
// over the points of convertZMapToXYZKernel (:570-593), whose invalid pixels
// hold z = 0: clamp(-z, min, max) mapped to [0, 1].
__global__ __launch_bounds__(256) void k_dbg_ref_depth(const float* __restrict__ z, int n, const float* __restrict__ c,
                                                       float* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float mn = c[MIND], mx = c[MAXD];
    const float zi = z[i];
    const float pz = (zi < -mn && zi > -mx) ? zi : 0.0f;
    float zval = -pz;
    if (zval < mn) zval = mn;
    if (zval > mx) zval = mx;
    out[i] = (zval - mn) / (mx - mn);
}

void launch_prepare_photo_reference(const Dev& d, hipStream_t s) {
    dim3 g16((d.W + 15) / 16, (d.H + 15) / 16);
    k_dbg_ref_depth<<<(d.W * d.H + 255) / 256, 256, 0, s>>>(d.ref.z, d.W * d.H, d.calib, d.dbg_ref);
    k_ref_features<<<g16, 256, 0, s>>>(d.ref.z, d.ref.gray[0], d.W, d.H, d.calib, d.gmag, d.ghist);
    k_select_count<<<512, 256, 0, s>>>(d.gmag, d.W * d.H, d.ghist, d.N, d.seg);
    k_select_scan<<<1, 512, 0, s>>>(d.seg);
    k_select_scatter<<<512, 256, 0, s>>>(d.gmag, d.W * d.H, d.N, d.seg, d.sel);
    k_ref_points<<<(d.N + 255) / 256, 256, 0, s>>>(d.ref.z, d.ref.gray[0], d.ref.gray[1], d.ref.gray[2], d.ref.gray[3],
                                                    d.W, d.H, d.L, d.N, d.calib, 1e-3f, d.sel, d.pts, d.jac[0],
                                                    d.jac[1], d.jac[2], d.jac[3]);
}

// Partial sums of the normal equations: FP64 per workgroup, term-major
// (part[k * PART_STRIDE + group]) so that the combining workgroup reads each
// term's partials contiguously. The system passes run at most SYS_GROUPS
// workgroups of SYS_THREADS: the combine reads partials that other XCDs have
// just written, and its time grows with their number of cache lines.

// Reduce the 27 per-thread FP64 sums of an NT-thread workgroup (per wave by
// halving exchanges, then the waves in order) into part[k * stride + group].
// The per-point products are FP32, their sums FP64: the FP32 oracle's
// accumulation, so that summation order leaves no trace at FP32 resolution.
template <int NT>
__device__ __forceinline__ void reduce27(const double* acc, double* part, int group, int stride = PART_STRIDE) {
    constexpr int NW = NT / 64;
    __shared__ double wsum[NW][27];
    const int wave = threadIdx.x >> 6, lane = lane_id();
    double v32[32];
#pragma unroll
    for (int k = 0; k < 32; k++) v32[k] = k < 27 ? acc[k] : 0.0;
    const double tot = wave_sum32d(v32);
    if ((lane & 1) == 0 && (lane >> 1) < 27) wsum[wave][lane >> 1] = tot;
    __syncthreads();
    if (threadIdx.x < 27) {
        const int k = threadIdx.x;
        double sum = wsum[0][k];
#pragma unroll
        for (int w = 1; w < NW; w++) sum += wsum[w][k];
        part[(size_t)k * stride + group] = sum;
    }
}
// ------------------------------------------------------------------ A10 + A11
// interpolateResidualKernel2 (warp.cu:269-332) with the level mapping of
// interpolateResidual2 (:1160-1218); the MAD histogram of discretizeResidual
// (histogram64.cu:41-44) is built in LDS.
template <int NT>
__device__ __forceinline__ void photo_res_body(const float4* __restrict__ pts, const float4* __restrict__ jac, int N,
                                               int layer, const float* __restrict__ img, int wl, int hl,
                                               const float* __restrict__ zcur, int W, int H,
                                               const float* __restrict__ c, const float* __restrict__ trel, int bins,
                                               float2* __restrict__ res, unsigned* __restrict__ madhist, int grp,
                                               int ngrp) {
    __shared__ unsigned h[256];
    if (threadIdx.x < 256) h[threadIdx.x] = 0u;
    float T[16];
#pragma unroll
    for (int k = 0; k < 16; k++) T[k] = trel[k];
    __syncthreads();
    float a = 1.0f / (float)(1 << layer);
    float b = 0.5f * (a - 1.0f);
    JacView jv = jac_view(const_cast<float4*>(jac), N);
    for (int i = grp * NT + threadIdx.x; i < N; i += ngrp * NT) {
        float ir, zW;
        photo_point(pts[i], jv.iref[i], T, a, b, F32Src{img}, wl, hl, F32Src{zcur}, W, H, c, ir, zW);
        res[i] = make_float2(ir, zW);
        int bn = (int)((float)(bins - 1) * fabsf(ir));
        if (bn >= 0 && bn < bins) atomicAdd(&h[bn], 1u);
    }
    __syncthreads();
    if (threadIdx.x < bins) {
        unsigned v = h[threadIdx.x];
        if (v) atomicAdd(&madhist[threadIdx.x], v);
    }
}
__global__ __launch_bounds__(256) void k_photo_res(const float4* __restrict__ pts, const float4* __restrict__ jac,
                                                   int N, int layer, const float* __restrict__ img, int wl, int hl,
                                                   const float* __restrict__ zcur, int W, int H,
                                                   const float* __restrict__ c, const float* __restrict__ trel,
                                                   int bins, float2* __restrict__ res, unsigned* __restrict__ madhist) {
    photo_res_body<256>(pts, jac, N, layer, img, wl, hl, zcur, W, H, c, trel, bins, res, madhist, blockIdx.x,
                        gridDim.x);
}

// The MAD histogram's median bin (generateWeights64Kernel, histogram64.cu:79-94):
// the first bin whose running count exceeds N / 2 (0 if none), by one wave:
// lane l holds bins 4l .. 4l+3, a shuffle scan gives the running counts.
__device__ __forceinline__ void med_bin(const unsigned* __restrict__ madhist, int bins, int N, int* med) {
    const int lane = threadIdx.x;
    const uint4 h4 = reinterpret_cast<const uint4*>(madhist)[lane];
    const int b0 = 4 * lane;
    const int c0 = b0 < bins ? (int)h4.x : 0, c1 = b0 + 1 < bins ? (int)h4.y : 0;
    const int c2 = b0 + 2 < bins ? (int)h4.z : 0, c3 = b0 + 3 < bins ? (int)h4.w : 0;
    const int tot = c0 + c1 + c2 + c3;
    int inc = tot;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(inc, o, 64);
        if (lane >= o) inc += v;
    }
    const int half = N / 2;
    int r = inc - tot;
    int first = 256;
    r += c0; if (first == 256 && r > half) first = b0;
    r += c1; if (first == 256 && r > half) first = b0 + 1;
    r += c2; if (first == 256 && r > half) first = b0 + 2;
    r += c3; if (first == 256 && r > half) first = b0 + 3;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) first = min(first, __shfl_xor(first, o, 64));
    if (lane == 0) *med = first < bins ? first : 0;
}
// Raw geometric partials (nraw groups, stride RAW_STRIDE) folded into nf
// groups: folded group f is the sum of raw groups f, f + nf, ... in order.
// One thread per term.
__device__ __forceinline__ void fold_partials(const double* __restrict__ raw, int nraw, double* __restrict__ dst, int nf,
                                              int f, int k) {
    if (f < nf && k < 27) {
        double sum = 0.0;
        for (int g = f; g < nraw; g += nf) sum += raw[(size_t)k * RAW_STRIDE + g];
        dst[(size_t)k * PART_STRIDE + f] = sum;
    }
}

// generateWeights64Kernel (histogram64.cu:79-94) + weightJacobianKernel
// (warp.cu:770-780) + the 27 dot products (warp.cu:1453-1487): per workgroup
// FP64 sums of the FP32 products (term-major in part); k_solve combines them.
template <int NT>
__device__ __forceinline__ void photo_sys_body(const float2* __restrict__ res, const float4* __restrict__ jac, int N,
                                               int bins, int robust, const unsigned* __restrict__ madhist,
                                               float* __restrict__ wts, double* part, int grp, int ngrp) {
    __shared__ int med;
    const int t = threadIdx.x;
    if (t < 64) med_bin(madhist, bins, N, &med);
    __syncthreads();
    float s = mad_scale(med, bins);
    double acc[27];
#pragma unroll
    for (int k = 0; k < 27; k++) acc[k] = 0.0;
    JacView jv = jac_view(const_cast<float4*>(jac), N);
    for (int i = grp * NT + t; i < N; i += ngrp * NT) {
        float2 rz = res[i];
        float r = rz.x;
        float w = robust_weight(r, rz.y, s, robust);
        float wr = r * w;
        wts[i] = w;
        float4 ja = jv.a[i];
        float2 jb = jv.b[i];
        float jw[6] = {ja.x * w, ja.y * w, ja.z * w, ja.w * w, jb.x * w, jb.y * w};
        int k = 0;
#pragma unroll
        for (int r0 = 0; r0 < 6; r0++)
#pragma unroll
            for (int c0 = r0; c0 < 6; c0++) acc[k++] += (double)(jw[r0] * jw[c0]);
#pragma unroll
        for (int r0 = 0; r0 < 6; r0++) acc[21 + r0] += (double)(jw[r0] * wr);
    }
    reduce27<NT>(acc, part, grp);
}
// The photometric system over gs workgroups; workgroups below nf also fold the
// geometric partials of the same iteration (wave 1, beside the median).
template <int NT>
__global__ __launch_bounds__(NT) void k_photo_sys(const float2* __restrict__ res, const float4* __restrict__ jac,
                                                  int N, int bins, int robust, const unsigned* __restrict__ madhist,
                                                  float* __restrict__ wts, double* part, int gs,
                                                  const double* __restrict__ raw, int nraw, double* fold, int nf) {
    if ((threadIdx.x >> 6) == 1) fold_partials(raw, nraw, fold, nf, blockIdx.x, threadIdx.x - 64);
    if ((int)blockIdx.x < gs) photo_sys_body<NT>(res, jac, N, bins, robust, madhist, wts, part, blockIdx.x, gs);
}

// The end of an iteration, one workgroup of 512 threads: the photometric
// partials (gp groups) into sys[0..27) and the geometric ones (gi groups) into
// sys[27..54) in FP64, the MAD histogram kept for inspection and cleared for
// the next iteration (madhist non-null), then, when solve is set, the FP64
// solve and the pose update on one wave (solve_update_wave) from an LDS copy
// of the system. A separate launch instead of the last workgroup of the system
// pass: electing that workgroup needs a device-scope fence in every workgroup,
// which writes back the L2.
// Wave w sums terms w, w + 8, ..., w + 48 of the 54; lane l adds groups
// l, l + 64, ... in order (all its loads issued first), then the lanes' sums
// meet in a fixed exchange tree: the same order for every call. No update
// from an empty geometric selection (nptr non-null with *nptr = 0).
constexpr int SOLVE_THREADS = 512;  // 8 waves, 2 per SIMD: room for the FP64 solve
constexpr int SOLVE_TERMS = 7;     // terms per wave (8 x 7 >= 54)
constexpr int COMBINE_LOADS = (PART_STRIDE + 63) / 64;  // groups per lane
__global__ __launch_bounds__(SOLVE_THREADS) void k_solve(const double* __restrict__ part, int gp,
                                                         const double* __restrict__ part2, int gi, unsigned* madhist,
                                                         const int* __restrict__ nptr, double* sys, float* trel,
                                                         SolveCfg cfg, int solve, int N1) {
    __shared__ double ws[SOLVE_WS];
    __shared__ double lsys[56];
    const int t = threadIdx.x, wave = t >> 6, lane = lane_id();
    const int n2 = nptr ? *nptr : 1;
    double v[SOLVE_TERMS][COMBINE_LOADS];
#pragma unroll
    for (int j = 0; j < SOLVE_TERMS; j++) {
        const int k = wave + 8 * j;
        const bool geo = k >= 27;
        const int groups = k >= 54 ? 0 : (geo ? gi : gp);
        const double* src = (geo ? part2 : part) + (size_t)(geo ? k - 27 : k) * PART_STRIDE;
#pragma unroll
        for (int i = 0; i < COMBINE_LOADS; i++) {
            const int g = lane + 64 * i;
            v[j][i] = g < groups ? src[g] : 0.0;
        }
    }
    unsigned mh = 0u;
    if (madhist && t < 256) mh = madhist[t];
#pragma unroll
    for (int j = 0; j < SOLVE_TERMS; j++) {
        const int k = wave + 8 * j;
        double acc = 0.0;
#pragma unroll
        for (int i = 0; i < COMBINE_LOADS; i++) acc += v[j][i];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
        const int groups = k >= 54 ? 0 : (k >= 27 ? gi : gp);
        if (lane == 0 && k < 54) {
            lsys[k] = acc;
            if (groups > 0) sys[k] = acc;
        }
    }
    if (madhist && t < 256) {
        madhist[256 + t] = mh;
        madhist[t] = 0u;
    }
    __syncthreads();
    if (solve && wave == 0 && n2 > 0) solve_update_wave(lsys, trel, cfg, N1, n2, ws);
}

int photo_groups(int N) {
    int g = N / 256;
    return g < 1 ? 1 : (g > 512 ? 512 : g);
}
// Workgroups of a system pass over n items (SYS_THREADS each, at most SYS_GROUPS).
static int sys_groups(int n) {
    int g = (n + SYS_THREADS - 1) / SYS_THREADS;
    return g < 1 ? 1 : (g > SYS_GROUPS ? SYS_GROUPS : g);
}
// The photometric system pass: 256-thread workgroups while SYS_GROUPS of them
// cover the points (one point per thread), else SYS_THREADS. Also folds nraw
// raw geometric partials into nf groups. Returns its number of groups.
static int launch_photo_sys(const Dev& d, int layer, const SolveCfg& c, const double* raw, int nraw, double* fold,
                            int nf, hipStream_t s) {
    const int g256 = (d.N + 255) / 256;
    if (g256 <= SYS_GROUPS) {
        const int gs = g256 < 1 ? 1 : g256;
        k_photo_sys<256><<<gs > nf ? gs : nf, 256, 0, s>>>(d.res, d.jac[layer], d.N, c.mad_bins, c.robust, d.madhist,
                                                           d.wts, d.part, gs, raw, nraw, fold, nf);
        return gs;
    }
    const int gs = sys_groups(d.N);
    k_photo_sys<SYS_THREADS><<<gs > nf ? gs : nf, SYS_THREADS, 0, s>>>(d.res, d.jac[layer], d.N, c.mad_bins, c.robust,
                                                                       d.madhist, d.wts, d.part, gs, raw, nraw, fold,
                                                                       nf);
    return gs;
}

void launch_photo_iteration(const Dev& d, int layer, const SolveCfg& c, int finish, hipStream_t s) {
    int G = photo_groups(d.N);
    k_photo_res<<<G, 256, 0, s>>>(d.pts, d.jac[layer], d.N, layer, d.cur.gray[layer], d.W >> layer, d.H >> layer,
                                  d.cur.z, d.W, d.H, d.calib, d.trel, c.mad_bins, d.res, d.madhist);
    const int Gs = launch_photo_sys(d, layer, c, nullptr, 0, nullptr, 0, s);
    k_solve<<<1, SOLVE_THREADS, 0, s>>>(d.part, Gs, nullptr, 0, d.madhist, nullptr, d.sys, d.trel, c,
                                        finish == FINISH_SOLVE ? 1 : 0, d.N);
}

// ------------------------------------------------------------------ A18 + A19
// convertZMapToXYZ6Kernel (rgb2gray.cu:595-624) + computeNormalsCudaKernel
// (d2z.cu:503-537): positions and normals recomputed from the depth map.
__global__ __launch_bounds__(256) void k_ref_cloud(const float* __restrict__ z, int W, int H, const float* __restrict__ c,
                                                   float4* __restrict__ pos, float4* __restrict__ nrm) {
    int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= W || y >= H) return;
    ref_cloud_pixel(F32Src{z}, W, H, x, y, c, pos, nrm);
}

// genPointsCloud6Kernel (d2z.cu:436-469): every z < 0 is valid.
__global__ __launch_bounds__(256) void k_cur_cloud(const float* __restrict__ z, int W, int H, const float* __restrict__ c,
                                                   float4* __restrict__ pos) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= W * H) return;
    cur_cloud_pixel(F32Src{z}, W, i, c, pos);
}

// downSampleCloud6Kernel (d2z.cu:557-588): masked 2x2 mean, normal renormalised.
__global__ __launch_bounds__(256) void k_cloud_down(const float4* __restrict__ sp, const float4* __restrict__ sn, int Wd,
                                                    int Hd, float4* __restrict__ dp, float4* __restrict__ dn) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= Wd * Hd) return;
    cloud_down_pixel(sp, sn, Wd, i, dp, dn);
}

void launch_ref_cloud(const Dev& d, hipStream_t s) {
    dim3 g16((d.W + 15) / 16, (d.H + 15) / 16);
    k_ref_cloud<<<g16, 256, 0, s>>>(d.ref.z, d.W, d.H, d.calib, d.rpos[0], d.rnrm[0]);
}

// xyz2DiffKernel (warp.cu:1955-1999) at level 0: at each selected pixel,
// min(25 |r|, 1) of its photometric residual under the current pose (1 when
// it leaves the view); other pixels keep their values.
__global__ __launch_bounds__(256) void k_dbg_err_photo(const int* __restrict__ sel, const float4* __restrict__ pts,
                                                       const float4* __restrict__ jac0, int N,
                                                       const float* __restrict__ img, const float* __restrict__ zcur,
                                                       int W, int H, const float* __restrict__ c,
                                                       const float* __restrict__ trel, float* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= N) return;
    float T[16];
#pragma unroll
    for (int k = 0; k < 16; k++) T[k] = trel[k];
    const JacView jv = jac_view(const_cast<float4*>(jac0), N);
    float ir, zW;
    photo_point(pts[i], jv.iref[i], T, 1.0f, 0.0f, F32Src{img}, W, H, F32Src{zcur}, W, H, c, ir, zW);
    const int v = sel[i];
    if (v >= 0 && v < W * H) out[v] = fminf(fabsf(ir) * 25.0f, 1.0f);
}
// xyz6DiffNormalizedKernel (rgb2gray.cu:143-171): per keyframe pixel of a
// level, min(|z' - z_cur| / 30, 1) against the current point at the nearest
// pixel its transform projects to (depth intrinsics KL), 0 where either is
// not selected or it leaves the view.
__global__ __launch_bounds__(256) void k_dbg_err_icp(const float4* __restrict__ rpos, const float4* __restrict__ cpos,
                                                     int w, int h, const float* __restrict__ c,
                                                     const float* __restrict__ trel, float a, float b,
                                                     float* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= w * h) return;
    const float4 r = rpos[i];
    float res = 0.0f;
    if (r.w != 0.0f) {
        float T[16];
#pragma unroll
        for (int k = 0; k < 16; k++) T[k] = trel[k];
        const float* K = c + KL;
        const f3 q3 = mul4(T, f3{r.x, r.y, r.z});
        const float qx = q3.x / q3.z, qy = q3.y / q3.z;
        const float u = a * (K[0] * qx + K[2]) + b, v = a * (K[4] * qy + K[5]) + b;
        if (u >= 0 && v >= 0 && u < w - 1 && v < h - 1) {
            const int off2 = (int)(u + 0.5f) + (int)(v + 0.5f) * w;
            const float4 q = cpos[off2];
            if (q.w > 0.0f) res = fminf(fabsf(q3.z - q.z) / 30.0f, 1.0f);
        }
    }
    out[i] = res;
}
void launch_error_image_photo(const Dev& d, hipStream_t s) {
    k_dbg_err_photo<<<(d.N + 255) / 256, 256, 0, s>>>(d.sel, d.pts, d.jac[0], d.N, d.cur.gray[0], d.cur.z, d.W, d.H,
                                                       d.calib, d.trel, d.dbg_err);
}
void launch_error_image_icp(const Dev& d, hipStream_t s) {
    for (int l = 0; l < d.L; l++) {
        const int w = d.W >> l, h = d.H >> l;
        const float a = 1.0f / (float)(1 << l), b = 0.5f * (a - 1.0f);
        k_dbg_err_icp<<<(w * h + 255) / 256, 256, 0, s>>>(d.rpos[l], d.cpos[l], w, h, d.calib, d.trel, a, b,
                                                           d.dbg_icp[l]);
    }
}

// genOrientedPoints6Kernel (d2z.cu:471-502): the keyframe cloud from ray-cast
// points and normals; a point is kept where z < 0 and |n|^2 > 1e-7, every
// other record is zero.
__global__ __launch_bounds__(256) void k_ref_oriented(const float4* __restrict__ xyz, const float4* __restrict__ nrm,
                                                      int n, float4* __restrict__ pos, float4* __restrict__ rn) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float4 v = xyz[i], q = nrm[i];
    const float len = q.x * q.x + q.y * q.y + q.z * q.z;
    const bool keep = v.z < 0 && len > 1e-7f;
    pos[i] = keep ? make_float4(v.x, v.y, v.z, 1.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    rn[i] = keep ? make_float4(q.x, q.y, q.z, 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}
void launch_ref_oriented(const Dev& d, const float* xyz, const float* nrm, hipStream_t s) {
    const int n = d.W * d.H;
    k_ref_oriented<<<(n + 255) / 256, 256, 0, s>>>((const float4*)xyz, (const float4*)nrm, n, d.rpos[0], d.rnrm[0]);
}
void launch_ref_cloud_pyramid(const Dev& d, hipStream_t s) {
    for (int l = 1; l < d.L; l++) {
        int n = (d.W >> l) * (d.H >> l);
        k_cloud_down<<<(n + 255) / 256, 256, 0, s>>>(d.rpos[l - 1], d.rnrm[l - 1], d.W >> l, d.H >> l, d.rpos[l],
                                                     d.rnrm[l]);
    }
}
void launch_cur_cloud(const Dev& d, hipStream_t s) {
    int n = d.W * d.H;
    k_cur_cloud<<<(n + 255) / 256, 256, 0, s>>>(d.cur.z, d.W, d.H, d.calib, d.cpos[0]);
    for (int l = 1; l < d.L; l++) {
        int m = (d.W >> l) * (d.H >> l);
        k_cloud_down<<<(m + 255) / 256, 256, 0, s>>>(d.cpos[l - 1], nullptr, d.W >> l, d.H >> l, d.cpos[l], nullptr);
    }
}

// ------------------------------------------------------------------ A20 + A21
// icpResidualKernel2 (rgb2gray.cu:223-260) over 1024-pixel row-major chunks;
// the valid flags are kept as wave ballots for the ordered compaction.
__global__ __launch_bounds__(1024) void k_icp_mask(const float4* __restrict__ rp, const float4* __restrict__ rn, int w,
                                                   int h, int layer, const float* __restrict__ trel,
                                                   const float* __restrict__ c, const float4* __restrict__ cp,
                                                   float* __restrict__ rfull, uint64_t* __restrict__ bits,
                                                   int* __restrict__ cnt) {
    __shared__ int wc[16];
    float T[16];
#pragma unroll
    for (int k = 0; k < 16; k++) T[k] = trel[k];
    int o = blockIdx.x * 1024 + threadIdx.x;
    bool valid = false;
    if (o < w * h) {
        const float a = 1.0f / (float)(1 << layer), b = 0.5f * (a - 1.0f);
        float r;
        valid = icp_mask_pixel(rp, rn, cp, w, h, o, T, c, a, b, r);
        if (valid) rfull[o] = r;
    }
    uint64_t bl = __ballot(valid);
    int wave = threadIdx.x >> 6;
    if (lane_id() == 0) {
        bits[(size_t)blockIdx.x * 16 + wave] = bl;
        wc[wave] = __popcll(bl);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int s = 0;
        for (int q = 0; q < 16; q++) s += wc[q];
        cnt[blockIdx.x] = s;
    }
}

// selectValidIndex + packIndex (rgb2gray.cu:827-850): ascending compaction and
// the squared residuals of the kept pixels.
__global__ __launch_bounds__(1024) void k_icp_compact(const float* __restrict__ rfull, const uint64_t* __restrict__ bits,
                                                      const int* __restrict__ cnt, int groups, int* __restrict__ sel,
                                                      float* __restrict__ r2, int* __restrict__ nout) {
    __shared__ int red[16];
    __shared__ int wpre[16];
    int g = blockIdx.x, t = threadIdx.x, wave = t >> 6, lane = lane_id();
    int v = 0;
    for (int q = t; q < g; q += 1024) v += cnt[q];
    v = (int)wave_sum((float)v);
    if (lane == 0) red[wave] = v;
    uint64_t word = bits[(size_t)g * 16 + wave];
    if (lane == 0) wpre[wave] = __popcll(word);
    __syncthreads();
    int prefix = 0;
    for (int q = 0; q < 16; q++) prefix += red[q];
    int wp = 0;
    for (int q = 0; q < wave; q++) wp += wpre[q];
    uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    if ((word >> lane) & 1ull) {
        int k = prefix + wp + __popcll(word & lt);
        int o = g * 1024 + t;
        float r = rfull[o];
        sel[k] = o;
        r2[k] = r * r;
    }
    if (g == groups - 1 && t == 0) {
        int tot = prefix;
        for (int q = 0; q < 16; q++) tot += wpre[q];
        *nout = tot;
    }
}

// generateStudentTWeights (rgb2gray.cu:470-568, 854-877), round k in 1..3: the
// weights of the previous round are recomputed from its variance, the products
// are truncated to int per element (one element per thread) and summed by the
// reference's 1024-wide tree into this round's block sums. The variance of the
// previous round, the reference's last-block combination of its block sums
// (int truncation per block sum, the same tree, clamped to at least n with
// more than one block), is recomputed by every block from those sums, and
// stored by block 0; k_student_finish combines the third round's.
__device__ __forceinline__ void tree1024(float* p, int t) {
    for (int s = 512; s >= 1; s >>= 1) {
        if (t < s) p[t] += p[t + s];
        __syncthreads();
    }
}
__device__ __forceinline__ float student_total(const float* __restrict__ bsum, int nb, int n, float* p, int t) {
    if (nb == 1) return bsum[0];
    int ss = 0;
    for (int q = t; q < nb; q += 1024) ss += bsum[q];
    p[t] = (float)ss;
    __syncthreads();
    tree1024(p, t);
    float v = p[0];
    const float mn = (float)n;
    if (v < mn) v = mn;
    __syncthreads();
    return v;
}
__global__ __launch_bounds__(1024) void k_student_round(const float* __restrict__ r2, const int* __restrict__ nptr,
                                                        int round, float* var, float* bsum, int stride) {
    __shared__ float p[1024];
    const int n = *nptr;
    const int t = threadIdx.x;
    const int nb = (n + 1023) / 1024;
    float prev = 0.0f;
    if (round > 1 && nb >= 1) {
        prev = student_total(bsum + (size_t)(round - 2) * stride, nb, n, p, t);
        if (blockIdx.x == 0 && t == 0) var[round - 2] = prev;
    }
    const int i = blockIdx.x * 1024 + t;
    int sum = 0;
    if (i < n) {
        float x = r2[i];
        float w = round == 1 ? 6 / (5 + x / 100.0f) : 6 / (5 + x / (prev / (float)n));
        sum += x * w;
    }
    p[t] = (float)sum;
    __syncthreads();
    tree1024(p, t);
    if (t == 0) bsum[(size_t)(round - 1) * stride + blockIdx.x] = p[0];
}
__global__ __launch_bounds__(1024) void k_student_finish(const int* __restrict__ nptr, float* var,
                                                         const float* __restrict__ bsum, int stride) {
    __shared__ float p[1024];
    const int n = *nptr;
    const int nb = (n + 1023) / 1024;
    if (nb < 1) return;
    const float v = student_total(bsum + 2 * (size_t)stride, nb, n, p, threadIdx.x);
    if (threadIdx.x == 0) var[2] = v;
}

// icpResidualIndexKernel2 (rgb2gray.cu:323-390) with the final clamped
// Student-t weight, and the 27 dot products of its system: per workgroup FP64
// sums of the FP32 products (term-major in part); k_solve combines them.
struct IcpSysArgs {
    const float4* rp;
    const float4* rn;
    const int* sel;
    const float* r2;
    const int* nptr;
    const float* var;
    int w, h, layer;
    const float* c;
    const float4* cp;
    const float* trel;
    double* part;
    int stride;  // of part
    float scale_in;
};
template <int NT>
__device__ __forceinline__ void icp_sys_body(const IcpSysArgs& g, int grp, int ngrp) {
    const float4* __restrict__ rp = g.rp;
    const float4* __restrict__ rn = g.rn;
    const int* __restrict__ sel = g.sel;
    const float* __restrict__ r2 = g.r2;
    const float* __restrict__ c = g.c;
    const float4* __restrict__ cp = g.cp;
    const float* __restrict__ trel = g.trel;
    const float* __restrict__ var = g.var;
    const int w = g.w, h = g.h, layer = g.layer;
    const float scale_in = g.scale_in;
    const int* nptr = g.nptr;
    int n = *nptr;
    float T[16];
#pragma unroll
    for (int k = 0; k < 16; k++) T[k] = trel[k];
    double acc[27];
#pragma unroll
    for (int k = 0; k < 27; k++) acc[k] = 0.0;
    float v3 = var[2];
    float a = 1.0f / (float)(1 << layer), b = 0.5f * (a - 1.0f);
    for (int i = grp * NT + threadIdx.x; i < n; i += ngrp * NT) {
        float res, J[6];
        icp_point_terms(rp, rn, cp, w, h, sel[i], r2[i], v3, n, T, c, a, b, scale_in, res, J);
        int k = 0;
#pragma unroll
        for (int r0 = 0; r0 < 6; r0++)
#pragma unroll
            for (int c0 = r0; c0 < 6; c0++) acc[k++] += (double)(J[r0] * J[c0]);
#pragma unroll
        for (int r0 = 0; r0 < 6; r0++) acc[21 + r0] += (double)(J[r0] * res);
    }
    reduce27<NT>(acc, g.part, grp, g.stride);
}
__global__ __launch_bounds__(SYS_THREADS) void k_icp_sys(IcpSysArgs g) {
    icp_sys_body<SYS_THREADS>(g, blockIdx.x, gridDim.x);
}
// A bi-objective iteration's photometric residuals (first gp workgroups) and
// its geometric system (the rest) in one launch: the geometric system does not
// depend on the residuals, so it leaves the path to the solve; each half works
// as its own kernel would.
__global__ __launch_bounds__(SYS_THREADS) void k_bi_res(const float4* __restrict__ pts,
                                                        const float4* __restrict__ jac, int N, int layer,
                                                        const float* __restrict__ img, int wl, int hl,
                                                        const float* __restrict__ zcur, int W, int H,
                                                        const float* __restrict__ c, int bins,
                                                        float2* __restrict__ res, unsigned* __restrict__ madhist,
                                                        int gp, IcpSysArgs g) {
    if ((int)blockIdx.x < gp)
        photo_res_body<SYS_THREADS>(pts, jac, N, layer, img, wl, hl, zcur, W, H, c, g.trel, bins, res, madhist,
                                    blockIdx.x, gp);
    else
        icp_sys_body<SYS_THREADS>(g, blockIdx.x - gp, gridDim.x - gp);
}

void launch_icp_select(const Dev& d, int layer, hipStream_t s) {
    int w = d.W >> layer, h = d.H >> layer, npix = w * h;
    int G = (npix + 1023) / 1024;
    uint64_t* words = d.icp_bits;
    k_icp_mask<<<G, 1024, 0, s>>>(d.rpos[layer], d.rnrm[layer], w, h, layer, d.trel, d.calib, d.cpos[layer], d.rfull,
                                  words, d.icp_cnt);
    k_icp_compact<<<G, 1024, 0, s>>>(d.rfull, words, d.icp_cnt, G, d.icp_sel[layer], d.icp_r2[layer],
                                     d.icp_n + layer);
    const int stride = (d.W * d.H + 1023) / 1024;
    for (int r = 1; r <= 3; r++)
        k_student_round<<<G, 1024, 0, s>>>(d.icp_r2[layer], d.icp_n + layer, r, d.icp_var + layer * 4, d.bsum, stride);
    k_student_finish<<<1, 1024, 0, s>>>(d.icp_n + layer, d.icp_var + layer * 4, d.bsum, stride);
}

static IcpSysArgs icp_sys_args(const Dev& d, int layer, const SolveCfg& c, double* part, int stride) {
    return IcpSysArgs{d.rpos[layer], d.rnrm[layer], d.icp_sel[layer], d.icp_r2[layer], d.icp_n + layer,
                      d.icp_var + layer * 4, d.W >> layer, d.H >> layer, layer, d.calib, d.cpos[layer], d.trel,
                      part, stride, c.scale_in};
}

void launch_icp_system(const Dev& d, int layer, const SolveCfg& c, int finish, hipStream_t s) {
    const int Gi = sys_groups((d.W >> layer) * (d.H >> layer));
    k_icp_sys<<<Gi, SYS_THREADS, 0, s>>>(icp_sys_args(d, layer, c, d.part2, PART_STRIDE));
    k_solve<<<1, SOLVE_THREADS, 0, s>>>(nullptr, 0, d.part2, Gi, nullptr, d.icp_n + layer, d.sys, d.trel, c,
                                        finish == FINISH_SOLVE ? 1 : 0, d.N);
}

// The geometric system runs beside the photometric residuals on up to
// RAW_STRIDE workgroups (every CU busy); the photometric system pass folds its
// partials to at most SYS_GROUPS for the combine.
void launch_biobjective_iteration(const Dev& d, int layer, const SolveCfg& c, int finish, bool select, hipStream_t s) {
    const int Gr = d.N > 0 ? (d.N + SYS_THREADS - 1) / SYS_THREADS : 1;  // residuals, one point per thread
    int Gi = ((d.W >> layer) * (d.H >> layer) + SYS_THREADS - 1) / SYS_THREADS;
    Gi = Gi < 1 ? 1 : (Gi > RAW_STRIDE ? RAW_STRIDE : Gi);
    const int nf = Gi < SYS_GROUPS ? Gi : SYS_GROUPS;
    if (select) launch_icp_select(d, layer, s);
    k_bi_res<<<Gr + Gi, SYS_THREADS, 0, s>>>(d.pts, d.jac[layer], d.N, layer, d.cur.gray[layer], d.W >> layer,
                                             d.H >> layer, d.cur.z, d.W, d.H, d.calib, c.mad_bins, d.res, d.madhist,
                                             Gr, icp_sys_args(d, layer, c, d.part3, RAW_STRIDE));
    const int Gs = launch_photo_sys(d, layer, c, d.part3, Gi, d.part2, nf, s);
    k_solve<<<1, SOLVE_THREADS, 0, s>>>(d.part, Gs, d.part2, nf, d.madhist, d.icp_n + layer, d.sys, d.trel, c,
                                        finish == FINISH_SOLVE ? 1 : 0, d.N);
}

// ------------------------------------------------------------------ A25
// Keyframe fusion: running weighted mean of the tracked frame's depth and gray
// into the keyframe through the relative pose (30 mm gate, weight cap 16).
__global__ __launch_bounds__(256) void k_fuse(float* __restrict__ kz, float* __restrict__ kg, float* __restrict__ kw,
                                              const float* __restrict__ fz, const float* __restrict__ fg, int W, int H,
                                              const float* __restrict__ trel, const float* __restrict__ c) {
    int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= W || y >= H) return;
    float T[16], iT[16];
#pragma unroll
    for (int k = 0; k < 16; k++) T[k] = trel[k];
    iT[0] = T[0]; iT[1] = T[4]; iT[2] = T[8];   iT[3] = -(T[0] * T[3] + T[4] * T[7] + T[8] * T[11]);
    iT[4] = T[1]; iT[5] = T[5]; iT[6] = T[9];   iT[7] = -(T[1] * T[3] + T[5] * T[7] + T[9] * T[11]);
    iT[8] = T[2]; iT[9] = T[6]; iT[10] = T[10]; iT[11] = -(T[2] * T[3] + T[6] * T[7] + T[10] * T[11]);
    iT[12] = 0.0f; iT[13] = 0.0f; iT[14] = 0.0f; iT[15] = 1.0f;
    size_t o = (size_t)x + (size_t)y * W;
    f3 p;
    if (!backproject(kz[o], x, y, c, p)) return;
    f3 q = mul4(T, p);
    const float* KLm = c + KL;
    float u = KLm[0] * (q.x / q.z) + KLm[2];
    float v = KLm[4] * (q.y / q.z) + KLm[5];
    if (!(u >= 0.0f && v >= 0.0f && u < (float)(W - 1) && v < (float)(H - 1))) return;
    int xi = (int)(u + 0.5f), yi = (int)(v + 0.5f);
    size_t fo = (size_t)xi + (size_t)yi * W;
    float zf = fz[fo];
    if (!(zf < -c[MIND] && zf > -c[MAXD])) return;
    float dz = zf - q.z;
    if (!(fabsf(dz) < 30.0f)) return;
    f3 f4;
    backproject(zf, xi, yi, c, f4);
    f3 pk = mul4(iT, f4);
    int xd = (int)u, yd = (int)v;
    float gobs = bilerp(fg, W, xd, yd, u - (float)xd, v - (float)yd);
    float Wold = kw[o];
    float Wn = Wold + 1.0f;
    kz[o] = (Wold * kz[o] + pk.z) / Wn;
    kg[o] = (Wold * kg[o] + gobs) / Wn;
    kw[o] = fminf(Wn, 16.0f);
}
void launch_fuse(const Dev& d, hipStream_t s) {
    dim3 g16((d.W + 15) / 16, (d.H + 15) / 16);
    k_fuse<<<g16, 256, 0, s>>>(d.ref.z, d.ref.gray[0], d.kf_w, d.cur.z, d.cur.gray[0], d.W, d.H, d.trel, d.calib);
}

// ------------------------------------------------------------------ A17
// PhotoTracker::applyIncrement (phototracker.cpp:514-531) on the device for
// sequences: pose[0..16) holds the host relative pose T (as float values),
// pose[16..32) the global pose.
__global__ void k_apply_increment(const float* trel, double* pose, double* out) {
    if (threadIdx.x != 0) return;
    double Tp[16], iTp[16], iT[16], Td[16], Tinc[16], G[16];
    for (int i = 0; i < 16; i++) { Tp[i] = pose[i]; iT[i] = trel[i]; G[i] = pose[16 + i]; }
    auto inv = [](const double* M, double* Mi) {
        Mi[0] = M[0]; Mi[1] = M[4]; Mi[2] = M[8];   Mi[3] = -(M[0] * M[3] + M[4] * M[7] + M[8] * M[11]);
        Mi[4] = M[1]; Mi[5] = M[5]; Mi[6] = M[9];   Mi[7] = -(M[1] * M[3] + M[5] * M[7] + M[9] * M[11]);
        Mi[8] = M[2]; Mi[9] = M[6]; Mi[10] = M[10]; Mi[11] = -(M[2] * M[3] + M[6] * M[7] + M[10] * M[11]);
        Mi[12] = 0; Mi[13] = 0; Mi[14] = 0; Mi[15] = 1;
    };
    auto mul = [](const double* A, const double* B, double* C) {
        double t[16];
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++)
                t[i * 4 + j] = A[i * 4 + 0] * B[0 * 4 + j] + A[i * 4 + 1] * B[1 * 4 + j] + A[i * 4 + 2] * B[2 * 4 + j] +
                               A[i * 4 + 3] * B[3 * 4 + j];
        for (int i = 0; i < 16; i++) C[i] = t[i];
    };
    inv(Tp, iTp);
    inv(iT, Td);
    mul(iTp, Td, Tinc);
    mul(G, Tinc, G);
    for (int i = 0; i < 16; i++) {
        pose[i] = (double)(float)Td[i];
        pose[16 + i] = G[i];
        if (out) out[i] = G[i];
    }
}
void launch_apply_increment(const Dev& d, double* pose_out, hipStream_t s) {
    k_apply_increment<<<1, 64, 0, s>>>(d.trel, d.pose, pose_out);
}

}  // namespace rt
