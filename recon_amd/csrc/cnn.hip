// Pose-proposal network: forward and backward kernels on the CDNA4 matrix
// cores (bfloat16 operands, FP32 accumulation), Adam, and the C ABI of
// include/rt_cnn.h.
//
// Every convolution is an implicit GEMM over NHWC bfloat16 tensors. The
// forward pass of the default network runs on v_mfma_f32_16x16x32_bf16
// (k_conv_fwd_mfma); other layer shapes and the backward passes use
// v_mfma_f32_16x16x16_bf16 tiles (64-lane wave, A: 16 rows x 4 k per lane
// group, B: 4 k x 16 columns, D: 4 rows x 16 columns per lane group):
//   forward      rows = output pixels (n, yo, xo), columns = Cout, k = (tap, Cin)
//   input grad   rows = input pixels (n, yi, xi),  columns = Cin,  k = (tap, Cout)
//   weight grad  rows = Cout, columns = (tap, Cin), k = output pixels, split
//                over workgroups and summed in a fixed order (deterministic)
// The head (global pooling, the 6-output dense layer, loss) is FP32.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/rt_cnn.h"
#include "cnn.hpp"
#include "rt_host.hpp"
#include "rt_stage.hpp"

namespace rt {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(bf16 v) { return __uint_as_float((uint32_t)v << 16); }
__device__ __forceinline__ bool bf_pos(bf16 v) { return !(v & 0x8000u) && v != 0; }

struct ConvGeom {
    int B, Hi, Wi, Ci, ci_shift, Ho, Wo, Co, co_shift, K, Kp;
};

__device__ __forceinline__ f4 mfma(s4 a, s4 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }

// ---------------------------------------------------------------- forward
// Block: 4 waves, 64 output pixels x 64 output channels; wave: 32 x 32.
__global__ __launch_bounds__(256) void k_conv_fwd(ConvGeom g, const bf16* __restrict__ X, const bf16* __restrict__ Wp,
                                                  const float* __restrict__ bias, bf16* __restrict__ Y) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r16 = lane & 15, kq = lane >> 4;
    const int M = g.B * g.Ho * g.Wo;
    const int row0 = blockIdx.x * 64 + (w & 1) * 32;
    const int col0 = blockIdx.y * 64 + (w >> 1) * 32;
    if (col0 >= g.Co || row0 >= M) return;
    const int HW = g.Ho * g.Wo;
    int nn[2], yo[2], xo[2];
    bool mv[2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
        int m = row0 + i * 16 + r16;
        mv[i] = m < M;
        int mm = mv[i] ? m : 0;
        nn[i] = mm / HW;
        int r = mm - nn[i] * HW;
        yo[i] = r / g.Wo;
        xo[i] = r - yo[i] * g.Wo;
    }
    f4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    const s4 zero = {0, 0, 0, 0};
    for (int k0 = kq * 4; k0 < g.Kp; k0 += 16) {
        const int tap = k0 >> g.ci_shift, ci = k0 & (g.Ci - 1);
        const int ky = tap / 3, kx = tap - 3 * (tap / 3);
        s4 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; i++) {
            int yi = 2 * yo[i] - 1 + ky, xi = 2 * xo[i] - 1 + kx;
            bool ok = mv[i] && tap < 9 && yi >= 0 && yi < g.Hi && xi >= 0 && xi < g.Wi;
            a[i] = ok ? *(const s4*)(X + (((size_t)nn[i] * g.Hi + yi) * g.Wi + xi) * g.Ci + ci) : zero;
        }
#pragma unroll
        for (int j = 0; j < 2; j++) {
            int c = col0 + j * 16 + r16;
            b[j] = c < g.Co ? *(const s4*)(Wp + (size_t)c * g.Kp + k0) : zero;
        }
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc[i][j] = mfma(a[i], b[j], acc[i][j]);
    }
#pragma unroll
    for (int j = 0; j < 2; j++) {
        int col = col0 + j * 16 + r16;
        if (col >= g.Co) continue;
        float bv = bias[col];
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                int m = row0 + i * 16 + kq * 4 + e;
                if (m < M) {
                    float v = acc[i][j][e] + bv;
                    Y[(size_t)m * g.Co + col] = f2bf(v > 0.0f ? v : 0.0f);
                }
            }
    }
}

// Forward with LDS tiles, for Cin a multiple of 32 and Cout a multiple of 64:
// block tile 128 output pixels x 64 output channels, k in steps of 32 (one tap,
// 32 consecutive input channels: 64 contiguous bytes per row). The im2col rows
// and the weights are staged in LDS once per block (rows padded to 72 bytes so
// that the 16 rows an MFMA operand reads hit distinct banks); each wave
// computes 32 pixels x 64 channels (2 x 4 MFMA tiles).
constexpr int FT_M = 128, FT_N = 64, FT_K = 32, FT_LD = FT_K + 4;  // LD in bf16 elements
__global__ __launch_bounds__(256) void k_conv_fwd_lds(ConvGeom g, const bf16* __restrict__ X,
                                                      const bf16* __restrict__ Wp, const float* __restrict__ bias,
                                                      bf16* __restrict__ Y) {
    __shared__ __attribute__((aligned(16))) bf16 As[FT_M * FT_LD];
    __shared__ __attribute__((aligned(16))) bf16 Bs[FT_N * FT_LD];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int r16 = lane & 15, kq = lane >> 4;
    const int M = g.B * g.Ho * g.Wo, HW = g.Ho * g.Wo;
    const int row0 = blockIdx.x * FT_M, col0 = blockIdx.y * FT_N;
    // this thread's two A rows (16-byte chunks: row = e / 4, chunk = e % 4)
    int an[2], ayo[2], axo[2], arow[2], achunk[2];
    bool av[2];
#pragma unroll
    for (int q = 0; q < 2; q++) {
        int e = t + q * 256;
        arow[q] = e >> 2;
        achunk[q] = e & 3;
        int m = row0 + arow[q];
        av[q] = m < M;
        int mm = av[q] ? m : 0;
        an[q] = mm / HW;
        int r = mm - an[q] * HW;
        ayo[q] = r / g.Wo;
        axo[q] = r - ayo[q] * g.Wo;
    }
    const int brow = t >> 2, bchunk = t & 3;
    f4 acc[2][4];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < g.K; k0 += FT_K) {
        const int tap = k0 >> g.ci_shift, ci0 = k0 & (g.Ci - 1);
        const int ky = tap / 3, kx = tap - 3 * (tap / 3);
#pragma unroll
        for (int q = 0; q < 2; q++) {
            int yi = 2 * ayo[q] - 1 + ky, xi = 2 * axo[q] - 1 + kx;
            uint4 v = {0u, 0u, 0u, 0u};
            if (av[q] && yi >= 0 && yi < g.Hi && xi >= 0 && xi < g.Wi)
                v = *(const uint4*)(X + (((size_t)an[q] * g.Hi + yi) * g.Wi + xi) * g.Ci + ci0 + 8 * achunk[q]);
            uint2* dst = (uint2*)(As + arow[q] * FT_LD + 8 * achunk[q]);
            dst[0] = make_uint2(v.x, v.y);
            dst[1] = make_uint2(v.z, v.w);
        }
        {
            uint4 v = *(const uint4*)(Wp + (size_t)(col0 + brow) * g.Kp + k0 + 8 * bchunk);
            uint2* dst = (uint2*)(Bs + brow * FT_LD + 8 * bchunk);
            dst[0] = make_uint2(v.x, v.y);
            dst[1] = make_uint2(v.z, v.w);
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < FT_K; kk += 16) {
            s4 a[2], b[4];
#pragma unroll
            for (int i = 0; i < 2; i++) a[i] = *(const s4*)(As + (w * 32 + i * 16 + r16) * FT_LD + kk + 4 * kq);
#pragma unroll
            for (int j = 0; j < 4; j++) b[j] = *(const s4*)(Bs + (j * 16 + r16) * FT_LD + kk + 4 * kq);
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[i][j] = mfma(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        int col = col0 + j * 16 + r16;
        float bv = bias[col];
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                int m = row0 + w * 32 + i * 16 + kq * 4 + e;
                if (m < M) {
                    float v = acc[i][j][e] + bv;
                    Y[(size_t)m * g.Co + col] = f2bf(v > 0.0f ? v : 0.0f);
                }
            }
    }
}

// Forward on gfx950's v_mfma_f32_16x16x32_bf16 (A: row lane % 16, eight
// consecutive k per lane group lane / 16; B the same k for column lane % 16;
// D: rows 4 * (lane / 16) + i, column lane % 16). Each wave owns a 16*TM x
// 16*TN tile; four waves (WGM x WGN) form the workgroup. Operands come
// straight from global memory as 16-byte loads (an 8-channel run of one tap
// of one input pixel, or 8 k of one weight row; two taps of 4 channels when
// Cin = 4), the next k-step's loads issued before the current step's MFMAs.
// The bias, ReLU and bfloat16 rounding are applied to the accumulators, and
// the tile leaves through LDS as 16-byte row chunks.
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ f4 mfma32(bf8 a, bf8 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
template <int TM, int TN, int WGM, int WGN, bool CI4>
__global__ __launch_bounds__(256) void k_conv_fwd_mfma(ConvGeom g, const bf16* __restrict__ X,
                                                       const bf16* __restrict__ Wp, const float* __restrict__ bias,
                                                       bf16* __restrict__ Y) {
    static_assert(WGM * WGN == 4, "four waves");
    constexpr int WM = 16 * TM, WN = 16 * TN, LDT = WN + 8;  // LDS row pitch (bf16)
    __shared__ __attribute__((aligned(16))) bf16 tile[4][WM * LDT];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r16 = lane & 15, kg = lane >> 4;
    const int M = g.B * g.Ho * g.Wo, HW = g.Ho * g.Wo;
    const int row0 = blockIdx.x * (WM * WGM) + (w % WGM) * WM;
    const int col0 = blockIdx.y * (WN * WGN) + (w / WGM) * WN;
    int yb[TM], xb[TM];
    size_t nb[TM];
    bool mv[TM];
#pragma unroll
    for (int t = 0; t < TM; t++) {
        const int m = row0 + 16 * t + r16;
        mv[t] = m < M;
        const int mm = mv[t] ? m : 0;
        const int n = mm / HW, r = mm - n * HW, yo = r / g.Wo, xo = r - yo * g.Wo;
        yb[t] = 2 * yo - 1;
        xb[t] = 2 * xo - 1;
        nb[t] = (size_t)n * g.Hi * g.Wi;
    }
    const bf16* wrow[TN];
#pragma unroll
    for (int j = 0; j < TN; j++) wrow[j] = Wp + (size_t)(col0 + 16 * j + r16) * g.Kp + 8 * kg;
    auto tap_ok = [&](int t, int tap, size_t& off) {
        const int ky = tap / 3, kx = tap - 3 * (tap / 3);
        const int yi = yb[t] + ky, xi = xb[t] + kx;
        off = (nb[t] + (size_t)yi * g.Wi + xi) * g.Ci;
        return mv[t] && tap < 9 && yi >= 0 && yi < g.Hi && xi >= 0 && xi < g.Wi;
    };
    auto load = [&](int s, bf8* a, bf8* b) {
        const int k = 32 * s + 8 * kg;
        if constexpr (CI4) {
            const int tap0 = k >> 2;
#pragma unroll
            for (int t = 0; t < TM; t++) {
                size_t o0, o1;
                const bool ok0 = tap_ok(t, tap0, o0), ok1 = tap_ok(t, tap0 + 1, o1);
                const uint2 v0 = ok0 ? *(const uint2*)(X + o0) : make_uint2(0u, 0u);
                const uint2 v1 = ok1 ? *(const uint2*)(X + o1) : make_uint2(0u, 0u);
                const uint4 v = make_uint4(v0.x, v0.y, v1.x, v1.y);
                a[t] = __builtin_bit_cast(bf8, v);
            }
        } else {
            const int tap = k >> g.ci_shift, ci = k & (g.Ci - 1);
#pragma unroll
            for (int t = 0; t < TM; t++) {
                size_t o;
                const bool ok = tap_ok(t, tap, o);
                const uint4 v = ok ? *(const uint4*)(X + o + ci) : make_uint4(0u, 0u, 0u, 0u);
                a[t] = __builtin_bit_cast(bf8, v);
            }
        }
#pragma unroll
        for (int j = 0; j < TN; j++) b[j] = __builtin_bit_cast(bf8, *(const uint4*)(wrow[j] + 32 * s));
    };
    f4 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    const int steps = g.Kp / 32;
    bf8 a[TM], b[TN];
    load(0, a, b);
    for (int s = 0; s < steps; s++) {
        bf8 an[TM], bn[TN];
        if (s + 1 < steps) load(s + 1, an, bn);
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
            for (int j = 0; j < TN; j++) acc[i][j] = mfma32(a[i], b[j], acc[i][j]);
#pragma unroll
        for (int i = 0; i < TM; i++) a[i] = an[i];
#pragma unroll
        for (int j = 0; j < TN; j++) b[j] = bn[j];
    }
    bf16* T = tile[w];
#pragma unroll
    for (int j = 0; j < TN; j++) {
        const float bv = bias[col0 + 16 * j + r16];
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const float v = acc[i][j][e] + bv;
                T[(16 * i + 4 * kg + e) * LDT + 16 * j + r16] = f2bf(v > 0.0f ? v : 0.0f);
            }
    }
    __syncthreads();
    constexpr int CPR = WN / 8;          // 16-byte chunks per row
    constexpr int RPI = 64 / CPR;        // rows per wave instruction
#pragma unroll
    for (int r = 0; r < WM; r += RPI) {
        const int rr = r + lane / CPR, c = lane % CPR;
        const int m = row0 + rr;
        const uint4 v = *(const uint4*)(T + rr * LDT + 8 * c);
        if (m < M) *(uint4*)(Y + (size_t)m * g.Co + col0 + 8 * c) = v;
    }
}

// Forward for Cin a multiple of 32 (the layers after the first): a workgroup
// tile of BM output pixels x BN output channels, k in steps of 32 (one tap, 32
// channels). Each step's A rows (64-byte runs of the input) and B rows (64
// bytes of weights) are staged in LDS by the whole workgroup, double-buffered:
// the next step's global loads are issued before this step's MFMAs and stored
// to the other buffer after them, one barrier per step. Waves (WGM x 4/WGM)
// take 16*TM x 16*TN sub-tiles with v_mfma_f32_16x16x32_bf16; the epilogue is
// the one of k_conv_fwd_mfma.
template <int BM, int BN, int WGM, int KS = 32>
__global__ __launch_bounds__(256) void k_conv_fwd_lds32(ConvGeom g, const bf16* __restrict__ X,
                                                        const bf16* __restrict__ Wp, const float* __restrict__ bias,
                                                        bf16* __restrict__ Y) {
    constexpr int WGN = 4 / WGM, WM = BM / WGM, WN = BN / WGN, TM = WM / 16, TN = WN / 16;
    constexpr int LD = KS + 8;  // LDS row pitch in bf16: KS k + 8
    constexpr int STAGE = (BM + BN) * LD;
    constexpr int CR = KS / 8;  // 16-byte chunks per row and stage
    constexpr int AC = BM * CR, BC = BN * CR;
    constexpr int AQ = (AC + 255) / 256, BQ = (BC + 255) / 256;
    constexpr int TLD = BN + 8;  // epilogue tile pitch
    static_assert(BM * TLD <= 2 * STAGE, "epilogue tile fits in the stage buffers");
    __shared__ __attribute__((aligned(16))) bf16 lds[2 * STAGE];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, r16 = lane & 15, kg = lane >> 4;
    const int M = g.B * g.Ho * g.Wo, HW = g.Ho * g.Wo;
    const int bm0 = blockIdx.x * BM, bn0 = blockIdx.y * BN;
    int arow[AQ], apart[AQ], ayb[AQ], axb[AQ];
    size_t anb[AQ];
    bool av[AQ], ae[AQ];
#pragma unroll
    for (int q = 0; q < AQ; q++) {
        const int e = t + 256 * q;
        ae[q] = e < AC;
        arow[q] = (e / CR) % BM;
        apart[q] = e % CR;
        const int m = bm0 + arow[q];
        av[q] = ae[q] && m < M;
        const int mm = av[q] ? m : 0;
        const int n = mm / HW, r = mm - n * HW, yo = r / g.Wo, xo = r - yo * g.Wo;
        ayb[q] = 2 * yo - 1;
        axb[q] = 2 * xo - 1;
        anb[q] = (size_t)n * g.Hi * g.Wi;
    }
    int brow[BQ], bpart[BQ];
    bool be[BQ];
    const bf16* bp[BQ];
#pragma unroll
    for (int q = 0; q < BQ; q++) {
        const int e = t + 256 * q;
        be[q] = e < BC;
        brow[q] = (e / CR) % BN;
        bpart[q] = e % CR;
        bp[q] = Wp + (size_t)(bn0 + brow[q]) * g.Kp + 8 * bpart[q];
    }
    uint4 ra[AQ], rb[BQ];
    auto gload = [&](int s) {
        const int k = KS * s;
        const int tap = k >> g.ci_shift, ci0 = k & (g.Ci - 1);
        const int ky = tap / 3, kx = tap - 3 * (tap / 3);
#pragma unroll
        for (int q = 0; q < AQ; q++) {
            const int yi = ayb[q] + ky, xi = axb[q] + kx;
            const bool ok = av[q] && yi >= 0 && yi < g.Hi && xi >= 0 && xi < g.Wi;
            ra[q] = ok ? *(const uint4*)(X + (anb[q] + (size_t)yi * g.Wi + xi) * g.Ci + ci0 + 8 * apart[q])
                       : make_uint4(0u, 0u, 0u, 0u);
        }
#pragma unroll
        for (int q = 0; q < BQ; q++) rb[q] = be[q] ? *(const uint4*)(bp[q] + k) : make_uint4(0u, 0u, 0u, 0u);
    };
    auto sstore = [&](int buf) {
        bf16* A = lds + buf * STAGE;
        bf16* Bs = A + BM * LD;
#pragma unroll
        for (int q = 0; q < AQ; q++)
            if (ae[q]) *(uint4*)(A + arow[q] * LD + 8 * apart[q]) = ra[q];
#pragma unroll
        for (int q = 0; q < BQ; q++)
            if (be[q]) *(uint4*)(Bs + brow[q] * LD + 8 * bpart[q]) = rb[q];
    };
    const int wm0 = (w % WGM) * WM, wn0 = (w / WGM) * WN;
    f4 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    const int steps = g.Kp / KS;
    gload(0);
    sstore(0);
    __syncthreads();
    for (int s = 0; s < steps; s++) {
        const int cur = s & 1;
        if (s + 1 < steps) gload(s + 1);
        const bf16* A = lds + cur * STAGE;
        const bf16* Bs = A + BM * LD;
#pragma unroll
        for (int kk = 0; kk < KS; kk += 32) {
            bf8 a[TM], b[TN];
#pragma unroll
            for (int i = 0; i < TM; i++) a[i] = *(const bf8*)(A + (wm0 + 16 * i + r16) * LD + kk + 8 * kg);
#pragma unroll
            for (int j = 0; j < TN; j++) b[j] = *(const bf8*)(Bs + (wn0 + 16 * j + r16) * LD + kk + 8 * kg);
#pragma unroll
            for (int i = 0; i < TM; i++)
#pragma unroll
                for (int j = 0; j < TN; j++) acc[i][j] = mfma32(a[i], b[j], acc[i][j]);
        }
        if (s + 1 < steps) sstore(cur ^ 1);
        __syncthreads();
    }
    bf16* T = lds;
#pragma unroll
    for (int j = 0; j < TN; j++) {
        const float bv = bias[bn0 + wn0 + 16 * j + r16];
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const float v = acc[i][j][e] + bv;
                T[(wm0 + 16 * i + 4 * kg + e) * TLD + wn0 + 16 * j + r16] = f2bf(v > 0.0f ? v : 0.0f);
            }
    }
    __syncthreads();
    constexpr int CPR = BN / 8;
#pragma unroll
    for (int e0 = 0; e0 < BM * CPR; e0 += 256) {
        const int e = e0 + t;
        if (e < BM * CPR) {
            const int rr = e / CPR, c = e - rr * CPR;
            const int m = bm0 + rr;
            if (m < M) *(uint4*)(Y + (size_t)m * g.Co + bn0 + 8 * c) = *(const uint4*)(T + rr * TLD + 8 * c);
        }
    }
}

// The first two forward layers fused, one workgroup (8 waves) per pair: the
// first layer (4 -> C1 channels, direct 8-byte tap loads as in
// k_conv_fwd_mfma) writes its H1 x W1 x C1 output to LDS with bias, ReLU and
// bfloat16 rounding; after a barrier the second layer (C1 = 32 -> C2) reads
// its A operands from that LDS map (one 16-byte run per lane and tap) and its
// weights from global memory. The intermediate never reaches HBM; the
// second layer's tile leaves through LDS as 16-byte row chunks. Two
// workgroups fit in a CU's LDS.
constexpr int F12_C1 = 32;
constexpr int F12_LDS = 30 * 40 * F12_C1;  // bf16 elements: the default network's first-layer map
template <int C2>
__global__ __launch_bounds__(512) void k_conv12_fused(ConvGeom g1, ConvGeom g2, const bf16* __restrict__ X,
                                                      const bf16* __restrict__ W1, const float* __restrict__ b1,
                                                      const bf16* __restrict__ W2, const float* __restrict__ b2,
                                                      bf16* __restrict__ Y) {
    constexpr int NW = 8, TN2 = C2 / 16, C2LD = C2 + 8;
    __shared__ __attribute__((aligned(16))) bf16 map[F12_LDS];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, r16 = lane & 15, kg = lane >> 4;
    const int pair = blockIdx.x;
    const int H1 = g1.Ho, Wd1 = g1.Wo, M1 = H1 * Wd1, H2 = g2.Ho, Wd2 = g2.Wo, M2 = H2 * Wd2;
    const bf16* Xp = X + (size_t)pair * g1.Hi * g1.Wi * 4;
    // ---- first layer: K = 36 in two steps of 32 (taps 8s + 2kg, 8s + 2kg + 1)
    bf8 bw[2][2];
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
        for (int s = 0; s < 2; s++)
            bw[j][s] = __builtin_bit_cast(bf8, *(const uint4*)(W1 + (size_t)(16 * j + r16) * g1.Kp + 32 * s + 8 * kg));
    float bias1[2];
#pragma unroll
    for (int j = 0; j < 2; j++) bias1[j] = b1[16 * j + r16];
    const int RT1 = (M1 + 15) / 16;
    constexpr int G1 = 4;  // row tiles whose loads are in flight together
    for (int rt0 = w; rt0 < RT1; rt0 += G1 * NW) {
        bf8 a[G1][2];
#pragma unroll
        for (int u = 0; u < G1; u++) {
            const int m = (rt0 + NW * u) * 16 + r16;
            const bool mv = rt0 + NW * u < RT1 && m < M1;
            const int yb = 2 * (m / Wd1) - 1, xb = 2 * (m % Wd1) - 1;
#pragma unroll
            for (int s = 0; s < 2; s++) {
                uint2 v[2];
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int tap = 8 * s + 2 * kg + h;
                    const int yi = yb + tap / 3, xi = xb + tap % 3;
                    const bool ok = mv && tap < 9 && yi >= 0 && yi < g1.Hi && xi >= 0 && xi < g1.Wi;
                    v[h] = ok ? *(const uint2*)(Xp + ((size_t)yi * g1.Wi + xi) * 4) : make_uint2(0u, 0u);
                }
                a[u][s] = __builtin_bit_cast(bf8, make_uint4(v[0].x, v[0].y, v[1].x, v[1].y));
            }
        }
#pragma unroll
        for (int u = 0; u < G1; u++) {
            const int rt = rt0 + NW * u;
            if (rt >= RT1) continue;
#pragma unroll
            for (int j = 0; j < 2; j++) {
                f4 acc = f4{0.f, 0.f, 0.f, 0.f};
                acc = mfma32(a[u][0], bw[j][0], acc);
                acc = mfma32(a[u][1], bw[j][1], acc);
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const int mr = rt * 16 + 4 * kg + e;
                    const float v = acc[e] + bias1[j];
                    if (mr < M1) map[mr * F12_C1 + 16 * j + r16] = f2bf(v > 0.0f ? v : 0.0f);
                }
            }
        }
    }
    __syncthreads();
    // ---- second layer: K = 9 taps x 32 channels, one tap per step
    const int RT2 = (M2 + 15) / 16;
    constexpr int MAXRT = 3;  // row tiles per wave: the default network has 19 over 8 waves
    f4 acc[MAXRT][TN2];
#pragma unroll
    for (int q = 0; q < MAXRT; q++)
#pragma unroll
        for (int j = 0; j < TN2; j++) acc[q][j] = f4{0.f, 0.f, 0.f, 0.f};
    int yb[MAXRT], xb[MAXRT];
    bool mv[MAXRT];
#pragma unroll
    for (int q = 0; q < MAXRT; q++) {
        const int m = (w + NW * q) * 16 + r16;
        mv[q] = (w + NW * q) < RT2 && m < M2;
        yb[q] = 2 * (m / Wd2) - 1;
        xb[q] = 2 * (m % Wd2) - 1;
    }
    const bool any1 = w + NW < RT2, any2 = w + 2 * NW < RT2;
    for (int s = 0; s < 9; s++) {
        const int ky = s / 3, kx = s % 3;
        bf8 b[TN2];
#pragma unroll
        for (int j = 0; j < TN2; j++)
            b[j] = __builtin_bit_cast(bf8, *(const uint4*)(W2 + (size_t)(16 * j + r16) * g2.Kp + 32 * s + 8 * kg));
#pragma unroll
        for (int q = 0; q < MAXRT; q++) {
            if ((q == 1 && !any1) || (q == 2 && !any2)) continue;
            const int yi = yb[q] + ky, xi = xb[q] + kx;
            const bool ok = mv[q] && yi >= 0 && yi < H1 && xi >= 0 && xi < Wd1;
            const uint4 v = ok ? *(const uint4*)(map + (yi * Wd1 + xi) * F12_C1 + 8 * kg) : make_uint4(0u, 0u, 0u, 0u);
            const bf8 a = __builtin_bit_cast(bf8, v);
#pragma unroll
            for (int j = 0; j < TN2; j++) acc[q][j] = mfma32(a, b[j], acc[q][j]);
        }
    }
    __syncthreads();  // the map is read by every wave; its LDS now holds the output tile
    bf16* T = map;
#pragma unroll
    for (int q = 0; q < MAXRT; q++) {
        const int rt = w + NW * q;
        if (rt >= RT2) continue;
#pragma unroll
        for (int j = 0; j < TN2; j++) {
            const float bv = b2[16 * j + r16];
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int mr = rt * 16 + 4 * kg + e;
                const float v = acc[q][j][e] + bv;
                if (mr < M2) T[mr * C2LD + 16 * j + r16] = f2bf(v > 0.0f ? v : 0.0f);
            }
        }
    }
    __syncthreads();
    constexpr int CPR = C2 / 8;
    bf16* Yp = Y + (size_t)pair * M2 * C2;
    for (int e = t; e < M2 * CPR; e += 512) {
        const int rr = e / CPR, c = e - rr * CPR;
        *(uint4*)(Yp + (size_t)rr * C2 + 8 * c) = *(const uint4*)(T + rr * C2LD + 8 * c);
    }
}

// ---------------------------------------------------------------- input gradient
// dX = conv^T(dZ, W), masked by the ReLU of the layer input (X > 0), stored as
// the next (earlier) layer's dZ. WT: [Cin][9][Cout].
__global__ __launch_bounds__(256) void k_conv_bwd_data(ConvGeom g, const bf16* __restrict__ dZ,
                                                       const bf16* __restrict__ WT, const bf16* __restrict__ X,
                                                       bf16* __restrict__ dX) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r16 = lane & 15, kq = lane >> 4;
    const int M = g.B * g.Hi * g.Wi;
    const int row0 = blockIdx.x * 64 + (w & 1) * 32;
    const int col0 = blockIdx.y * 64 + (w >> 1) * 32;
    if (col0 >= g.Ci || row0 >= M) return;
    const int HW = g.Hi * g.Wi;
    int nn[2], yi[2], xi[2];
    bool mv[2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
        int m = row0 + i * 16 + r16;
        mv[i] = m < M;
        int mm = mv[i] ? m : 0;
        nn[i] = mm / HW;
        int r = mm - nn[i] * HW;
        yi[i] = r / g.Wi;
        xi[i] = r - yi[i] * g.Wi;
    }
    f4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    const s4 zero = {0, 0, 0, 0};
    const int Kd = 9 * g.Co;
    for (int k0 = kq * 4; k0 < Kd; k0 += 16) {
        const int tap = k0 >> g.co_shift, co = k0 & (g.Co - 1);
        const int ky = tap / 3, kx = tap - 3 * (tap / 3);
        s4 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; i++) {
            int ty = yi[i] + 1 - ky, tx = xi[i] + 1 - kx;
            int yo = ty >> 1, xo = tx >> 1;
            bool ok = mv[i] && ty >= 0 && tx >= 0 && !(ty & 1) && !(tx & 1) && yo < g.Ho && xo < g.Wo;
            a[i] = ok ? *(const s4*)(dZ + (((size_t)nn[i] * g.Ho + yo) * g.Wo + xo) * g.Co + co) : zero;
        }
#pragma unroll
        for (int j = 0; j < 2; j++) {
            int c = col0 + j * 16 + r16;
            b[j] = c < g.Ci ? *(const s4*)(WT + ((size_t)c * 9 + tap) * g.Co + co) : zero;
        }
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc[i][j] = mfma(a[i], b[j], acc[i][j]);
    }
#pragma unroll
    for (int j = 0; j < 2; j++) {
        int col = col0 + j * 16 + r16;
        if (col >= g.Ci) continue;
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                int m = row0 + i * 16 + kq * 4 + e;
                if (m < M) {
                    size_t o = (size_t)m * g.Ci + col;
                    dX[o] = f2bf(bf_pos(X[o]) ? acc[i][j][e] : 0.0f);
                }
            }
    }
}

// ---------------------------------------------------------------- weight gradient
// ws[s][co][k] = sum over output pixels m of chunk s of dZ[m][co] * im2col(X)[m][k].
// One wave per block: 32 output channels x 32 k.
__global__ __launch_bounds__(64) void k_conv_bwd_w(ConvGeom g, const bf16* __restrict__ dZ, const bf16* __restrict__ X,
                                                   float* __restrict__ ws, int chunk) {
    const int lane = threadIdx.x;
    const int r16 = lane & 15, kq = lane >> 4;
    const int M = g.B * g.Ho * g.Wo, HW = g.Ho * g.Wo;
    const int co0 = blockIdx.x * 32, kc0 = blockIdx.y * 32, s = blockIdx.z;
    const int mbeg = s * chunk, mend = min(M, mbeg + chunk);
    int tapj[2], cij[2];
    bool kv[2];
#pragma unroll
    for (int j = 0; j < 2; j++) {
        int k = kc0 + j * 16 + r16;
        tapj[j] = k >> g.ci_shift;
        cij[j] = k & (g.Ci - 1);
        kv[j] = k < g.K;
    }
    f4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    for (int m0 = mbeg; m0 < mend; m0 += 16) {
        s4 a[2], b[2];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            int m = m0 + kq * 4 + e;
            bool mok = m < mend;
            int mm = mok ? m : mbeg;
            int n = mm / HW, r = mm - (mm / HW) * HW;
            int yo = r / g.Wo, xo = r - (r / g.Wo) * g.Wo;
#pragma unroll
            for (int i = 0; i < 2; i++) {
                int co = co0 + i * 16 + r16;
                a[i][e] = (mok && co < g.Co) ? (short)dZ[(size_t)mm * g.Co + co] : (short)0;
            }
#pragma unroll
            for (int j = 0; j < 2; j++) {
                int tap = tapj[j];
                int ky = tap / 3, kx = tap - 3 * (tap / 3);
                int yi = 2 * yo - 1 + ky, xi = 2 * xo - 1 + kx;
                bool ok = mok && kv[j] && yi >= 0 && yi < g.Hi && xi >= 0 && xi < g.Wi;
                b[j][e] = ok ? (short)X[(((size_t)n * g.Hi + yi) * g.Wi + xi) * g.Ci + cij[j]] : (short)0;
            }
        }
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc[i][j] = mfma(a[i], b[j], acc[i][j]);
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            int k = kc0 + j * 16 + r16;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                int co = co0 + i * 16 + kq * 4 + e;
                if (co < g.Co && k < g.Kp) ws[((size_t)s * g.Co + co) * g.Kp + k] = acc[i][j][e];
            }
        }
}

// grad_w[co][k] = sum_s ws[s][co][k] (k < K), in order of s
__global__ void k_reduce_w(const float* __restrict__ ws, int S, int Co, int Kp, int K, float* __restrict__ gw) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= Co * K) return;
    int co = idx / K, k = idx - (idx / K) * K;
    float s = 0.0f;
    for (int q = 0; q < S; q++) s += ws[((size_t)q * Co + co) * Kp + k];
    gw[idx] = s;
}

// per-chunk column sums of dZ (bias gradient), then their ordered sum
__global__ void k_colsum(const bf16* __restrict__ dZ, int M, int C, int chunk, float* __restrict__ ws) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    int s = blockIdx.y;
    if (c >= C) return;
    int mbeg = s * chunk, mend = min(M, mbeg + chunk);
    float acc = 0.0f;
    for (int m = mbeg; m < mend; m++) acc += bf2f(dZ[(size_t)m * C + c]);
    ws[(size_t)s * C + c] = acc;
}
__global__ void k_reduce_rows(const float* __restrict__ ws, int S, int width, float* __restrict__ out) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= width) return;
    float s = 0.0f;
    for (int q = 0; q < S; q++) s += ws[(size_t)q * width + k];
    out[k] = s;
}

// ---------------------------------------------------------------- head
// One block per sample: mean pool, dense 6 outputs; with targets also the loss,
// the dense-layer gradient contributions and dZ of the last convolution.
// 1 / max(1, sum of the sample weights), in order
__global__ void k_weight_norm(const float* __restrict__ wn, int n, float* __restrict__ inv_norm) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float s = 0.0f;
    for (int i = 0; i < n; i++) s += wn[i];
    *inv_norm = 1.0f / fmaxf(1.0f, s);
}

__global__ __launch_bounds__(256) void k_head(int P6, int C6, const bf16* __restrict__ A6, const float* __restrict__ Wfc,
                                              const float* __restrict__ bfc, const float* __restrict__ y,
                                              const float* __restrict__ wn, const float* __restrict__ inv_norm_p,
                                              float* __restrict__ out,
                                              bf16* __restrict__ dZ6, float* __restrict__ ws_fc, float* __restrict__ ws_loss) {
    extern __shared__ float sh[];  // h[C6], dh[C6]
    __shared__ float red[4][6];
    __shared__ float o6[6], d6[6];
    const int n = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const float inv_norm = y ? *inv_norm_p : 0.0f;
    float* h = sh;
    for (int c = t; c < C6; c += 256) {
        float s = 0.0f;
        for (int p = 0; p < P6; p++) s += bf2f(A6[((size_t)n * P6 + p) * C6 + c]);
        h[c] = s / (float)P6;
    }
    __syncthreads();
    for (int j = 0; j < 6; j++) {
        float s = 0.0f;
        for (int c = t; c < C6; c += 256) s += Wfc[j * C6 + c] * h[c];
        for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
        if (lane == 0) red[wv][j] = s;
    }
    __syncthreads();
    if (t < 6) {
        float v = bfc[t] + ((red[0][t] + red[1][t]) + (red[2][t] + red[3][t]));
        o6[t] = v;
        out[n * 6 + t] = v;
        if (y) d6[t] = 2.0f * wn[n] * (v - y[n * 6 + t]) * inv_norm;
    }
    __syncthreads();
    if (!y) return;
    if (t == 0) {
        float l = 0.0f;
        for (int j = 0; j < 6; j++) { float e = o6[j] - y[n * 6 + j]; l += e * e; }
        ws_loss[n] = wn[n] * l * inv_norm;
    }
    const int width = 6 * C6 + 6;
    float* dh = sh + C6;
    for (int c = t; c < C6; c += 256) {
        float s = 0.0f;
        for (int j = 0; j < 6; j++) {
            ws_fc[(size_t)n * width + j * C6 + c] = d6[j] * h[c];
            s += Wfc[j * C6 + c] * d6[j];
        }
        dh[c] = s;
    }
    if (t < 6) ws_fc[(size_t)n * width + 6 * C6 + t] = d6[t];
    __syncthreads();
    for (int q = t; q < P6 * C6; q += 256) {
        int c = q % C6;
        size_t o = (size_t)n * P6 * C6 + q;
        dZ6[o] = f2bf(bf_pos(A6[o]) ? dh[c] / (float)P6 : 0.0f);
    }
}

// ---------------------------------------------------------------- parameters
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       size_t n, float lr, float b1, float b2, float c1, float c2, float eps) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    float mi = b1 * m[i] + (1.0f - b1) * gi;
    float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi; v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}
// bfloat16 copies of one layer's weights: Wp[co][Kp] (zero padded) and WT[ci][tap][co]
__global__ void k_pack(const float* __restrict__ W, int Co, int Ci, int Kp, bf16* __restrict__ Wp, bf16* __restrict__ WT) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= Co * Kp) return;
    int co = idx / Kp, k = idx - (idx / Kp) * Kp;
    int K = 9 * Ci;
    bf16 v = k < K ? f2bf(W[(size_t)co * K + k]) : (bf16)0;
    Wp[idx] = v;
    if (WT && k < K) {
        int tap = k / Ci, ci = k - (k / Ci) * Ci;
        WT[((size_t)ci * 9 + tap) * Co + co] = v;
    }
}

// ---------------------------------------------------------------- inputs and proposals
// The network's inputs, one cell of bw x bh = (W / IW) x (H / IH) pixels per
// thread and per frame (cell_features; the inverse depths by the hardware
// reciprocal: within 1 ulp, far below the bfloat16 rounding of the result).
// The inverse depths are summed per column over the rows in order; the
// column sums are merged pairwise by a stack (two sums of equal size merge,
// left + right; what remains at the end is added right to left), which for
// a power-of-two bw is the balanced tree ((c0 + c1) + (c2 + c3)) + ...
__global__ __launch_bounds__(256) void k_inputs(const FrameSrc* __restrict__ src, int IW, int IH, bf16* __restrict__ x) {
    const int n = blockIdx.y;
    const int cell = blockIdx.x * 256 + threadIdx.x;
    if (cell >= IW * IH) return;
    const FrameSrc s = src[n];
    const int bw = s.W / IW, bh = s.H / IH;
    const int cx = cell % IW, cy = cell / IW;
    bf16* o = x + ((size_t)n * IW * IH + cell) * 4;
    for (int f = 0; f < 2; f++) {
        const uint8_t* g8 = s.g8[f];
        const uint16_t* d16 = s.d16[f];
        const size_t row0 = (size_t)(cy * bh) * s.W + (size_t)cx * bw;
        if (bw == 8 && bh == 8 && !(((uintptr_t)(g8 + row0)) & 7) && !(((uintptr_t)(d16 + row0)) & 15) && !(s.W & 7)) {
            cell8_features(g8, d16, s.W, cx * 8, cy * 8, o + 2 * f);
            continue;
        }
        unsigned gs = 0, nv = 0;
        float sv[16];  // bw < 65536 columns: at most 16 pending sums
        int sl[16], sp = 0;
        for (int xx = 0; xx < bw; xx++) {
            float v = 0.0f;
            for (int yy = 0; yy < bh; yy++) {
                const size_t i = row0 + (size_t)yy * s.W + xx;
                gs += g8[i];
                unsigned d = d16[i];
                if (d) { v += inv_depth_m(d); nv++; }
            }
            int l = 0;
            while (sp > 0 && sl[sp - 1] == l) { v = sv[--sp] + v; l++; }
            sv[sp] = v; sl[sp] = l; sp++;
        }
        float ids = sv[sp - 1];
        for (int k = sp - 2; k >= 0; k--) ids = sv[k] + ids;
        cell_features(gs, ids, nv, bw * bh, o + 2 * f);
    }
}

// Training-set symmetries of a pinhole camera with a centred principal point.
// flags bit 0: mirror columns (camera x -> -x), bit 1: mirror rows (y -> -y),
// bit 2: swap reference and current frame (target = that of the inverse pose,
// y_inv). A mirror S maps Trel to S Trel S: rotation vector -S w, translation S t.
__global__ __launch_bounds__(256) void k_augment(const bf16* __restrict__ x, const float* __restrict__ y,
                                                 const float* __restrict__ y_inv, const int* __restrict__ flags, int IW,
                                                 int IH, bf16* __restrict__ xo, float* __restrict__ yo) {
    const int n = blockIdx.y;
    const int cell = blockIdx.x * 256 + threadIdx.x;
    const int f = flags[n];
    if (cell < IW * IH) {
        int cx = cell % IW, cy = cell / IW;
        int sx = (f & 1) ? IW - 1 - cx : cx, sy = (f & 2) ? IH - 1 - cy : cy;
        const bf16* src = x + ((size_t)n * IW * IH + (size_t)sy * IW + sx) * 4;
        bf16* dst = xo + ((size_t)n * IW * IH + cell) * 4;
        if (f & 4) { dst[0] = src[2]; dst[1] = src[3]; dst[2] = src[0]; dst[3] = src[1]; }
        else { dst[0] = src[0]; dst[1] = src[1]; dst[2] = src[2]; dst[3] = src[3]; }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const float* b = (f & 4) ? y_inv + 6 * n : y + 6 * n;
        float v[6] = {b[0], b[1], b[2], b[3], b[4], b[5]};
        if (f & 1) { v[1] = -v[1]; v[2] = -v[2]; v[3] = -v[3]; }
        if (f & 2) { v[0] = -v[0]; v[2] = -v[2]; v[4] = -v[4]; }
        for (int k = 0; k < 6; k++) yo[6 * n + k] = v[k];
    }
}

// Trel = [R(w) | t] from the scaled outputs (Rodrigues), FP64, stored FP32
__global__ void k_outputs_to_trel(const float* __restrict__ out, int n, float rot_scale_rad, float trans_scale,
                                  float* __restrict__ trel) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double w[3], t[3];
    for (int k = 0; k < 3; k++) { w[k] = (double)out[6 * i + k] * rot_scale_rad; t[k] = (double)out[6 * i + 3 + k] * trans_scale; }
    double th = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    double A = th > 1e-12 ? sin(th) / th : 1.0;
    double B = th > 1e-12 ? (1.0 - cos(th)) / (th * th) : 0.5;
    double K[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
    double K2[9];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            double s = 0.0;
            for (int q = 0; q < 3; q++) s += K[3 * r + q] * K[3 * q + c];
            K2[3 * r + c] = s;
        }
    float* T = trel + 16 * (size_t)i;
    for (int r = 0; r < 3; r++) {
        for (int c = 0; c < 3; c++) T[4 * r + c] = (float)((r == c ? 1.0 : 0.0) + A * K[3 * r + c] + B * K2[3 * r + c]);
        T[4 * r + 3] = (float)t[r];
    }
    T[12] = 0.0f; T[13] = 0.0f; T[14] = 0.0f; T[15] = 1.0f;
}

}  // namespace rt

using namespace rt;

// ---------------------------------------------------------------- host object
struct rt_cnn {
    int device = 0, max_batch = 0;
    rt_cnn_config cfg;
    hipStream_t stream = nullptr;
    rt::HostStage st_src, st_flags;  // pinned staging of the frame table and the augmentation flags
    int H[7], W[7];
    size_t w_off[6], b_off[6], fc_w, fc_b, nparams;
    int K[6], Kp[6];
    std::vector<void*> allocs;
    float *params = nullptr, *grads = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    bf16* Wp[6] = {};
    bf16* WT[6] = {};
    bf16* act[7] = {};          // act[0] is the caller's input during a call
    bf16 *dz_a = nullptr, *dz_b = nullptr;
    bf16* xin = nullptr;        // inputs built by rt_batch_propose
    float *ws = nullptr, *ws_b = nullptr, *ws_fc = nullptr, *ws_loss = nullptr, *out = nullptr, *loss = nullptr;
    size_t ws_floats = 0, ws_b_floats = 0;
    FrameSrc* src_dev = nullptr;
    int* flags_dev = nullptr;
    float* trel_tmp = nullptr;
    int step = 0;
    bool packed = false;
};

static int ilog2(int v) {
    int s = 0;
    while ((1 << s) < v) s++;
    return (1 << s) == v ? s : -1;
}

extern "C" void rt_cnn_default_config(rt_cnn_config* c) {
    if (!c) return;
    c->in_width = 80; c->in_height = 60;
    const int ch[7] = {4, 32, 64, 128, 128, 256, 256};
    for (int l = 0; l < 7; l++) c->channels[l] = ch[l];
    c->rot_scale_deg = 10.0f;
    c->trans_scale_mm = 150.0f;
}

static ConvGeom geom(const rt_cnn* net, int l, int B) {
    ConvGeom g;
    g.B = B; g.Hi = net->H[l]; g.Wi = net->W[l]; g.Ci = net->cfg.channels[l]; g.ci_shift = ilog2(g.Ci);
    g.Ho = net->H[l + 1]; g.Wo = net->W[l + 1]; g.Co = net->cfg.channels[l + 1]; g.co_shift = ilog2(g.Co);
    g.K = net->K[l]; g.Kp = net->Kp[l];
    return g;
}

static size_t act_elems(const rt_cnn* net, int l, int B) {
    return (size_t)B * net->H[l] * net->W[l] * net->cfg.channels[l];
}

// weight-gradient split: chunks of output pixels (multiple of 16) so that the
// grid holds about 2048 waves
static void wsplit(const rt_cnn* net, int l, int B, int* S, int* chunk) {
    int M = B * net->H[l + 1] * net->W[l + 1];
    int tiles = ((net->cfg.channels[l + 1] + 31) / 32) * (net->Kp[l] / 32);
    int s = std::max(1, std::min(2048 / std::max(tiles, 1), (M + 15) / 16));
    int c = ((M + s - 1) / s + 15) / 16 * 16;
    *chunk = c;
    *S = (M + c - 1) / c;
}

extern "C" int rt_cnn_create(int device, const rt_cnn_config* c, int max_batch, rt_cnn** out) {
    if (!out || max_batch < 1) return set_error(RT_ERR_ARG, "rt_cnn_create: bad argument");
    *out = nullptr;
    rt_cnn_config cfg;
    if (c) cfg = *c; else rt_cnn_default_config(&cfg);
    if (cfg.in_width < 2 || cfg.in_height < 2) return set_error(RT_ERR_ARG, "rt_cnn_create: input too small");
    if (cfg.channels[0] != 4) return set_error(RT_ERR_ARG, "rt_cnn_create: the input has 4 channels");
    for (int l = 1; l < 7; l++)
        if (ilog2(cfg.channels[l]) < 4 || cfg.channels[l] > 1024)
            return set_error(RT_ERR_ARG, "rt_cnn_create: conv channels must be powers of two in 16..1024");
    if (!(cfg.rot_scale_deg > 0) || !(cfg.trans_scale_mm > 0)) return set_error(RT_ERR_ARG, "rt_cnn_create: bad output scales");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return set_error(RT_ERR_ARG, "no such device");
    if (hipSetDevice(device) != hipSuccess) return set_error(RT_ERR_HIP, "hipSetDevice");
    rt_cnn* net = new rt_cnn();
    net->device = device; net->max_batch = max_batch; net->cfg = cfg;
    net->H[0] = cfg.in_height; net->W[0] = cfg.in_width;
    for (int l = 1; l < 7; l++) { net->H[l] = (net->H[l - 1] + 1) / 2; net->W[l] = (net->W[l - 1] + 1) / 2; }
    size_t o = 0;
    for (int l = 0; l < 6; l++) {
        net->K[l] = 9 * cfg.channels[l];
        net->Kp[l] = (net->K[l] + 31) / 32 * 32;
        net->w_off[l] = o; o += (size_t)cfg.channels[l + 1] * net->K[l];
        net->b_off[l] = o; o += (size_t)cfg.channels[l + 1];
    }
    net->fc_w = o; o += 6 * (size_t)cfg.channels[6];
    net->fc_b = o; o += 6;
    net->nparams = o;
    // workspace sizes
    size_t wsf = 0, wsb = 0, maxact = 0;
    for (int l = 0; l < 6; l++) {
        // wsplit never uses more than 2048 / tiles splits, whatever the batch
        int tiles = ((cfg.channels[l + 1] + 31) / 32) * (net->Kp[l] / 32);
        int S = std::max(1, 2048 / std::max(tiles, 1));
        wsf = std::max(wsf, (size_t)S * cfg.channels[l + 1] * net->Kp[l]);
        int M = max_batch * net->H[l + 1] * net->W[l + 1];
        int Sb = std::min(1024, (M + 255) / 256);
        wsb = std::max(wsb, (size_t)Sb * cfg.channels[l + 1]);
    }
    for (int l = 1; l < 7; l++) maxact = std::max(maxact, act_elems(net, l, max_batch));
    net->ws_floats = wsf; net->ws_b_floats = wsb;
    bool ok = hipStreamCreateWithFlags(&net->stream, hipStreamNonBlocking) == hipSuccess;
    auto al = [&](void** p, size_t bytes) {
        if (!ok) return;
        if (hipMalloc(p, std::max(bytes, (size_t)256)) != hipSuccess) { ok = false; return; }
        net->allocs.push_back(*p);
        if (hipMemset(*p, 0, std::max(bytes, (size_t)256)) != hipSuccess) ok = false;
    };
    al((void**)&net->params, 4 * o); al((void**)&net->grads, 4 * o);
    al((void**)&net->adam_m, 4 * o); al((void**)&net->adam_v, 4 * o);
    for (int l = 0; l < 6; l++) {
        al((void**)&net->Wp[l], 2 * (size_t)cfg.channels[l + 1] * net->Kp[l]);
        if (l > 0) al((void**)&net->WT[l], 2 * (size_t)cfg.channels[l + 1] * net->K[l]);
    }
    for (int l = 1; l < 7; l++) al((void**)&net->act[l], 2 * act_elems(net, l, max_batch));
    al((void**)&net->dz_a, 2 * maxact); al((void**)&net->dz_b, 2 * maxact);
    al((void**)&net->xin, 2 * act_elems(net, 0, max_batch));
    al((void**)&net->ws, 4 * wsf); al((void**)&net->ws_b, 4 * wsb);
    al((void**)&net->ws_fc, 4 * (size_t)max_batch * (6 * cfg.channels[6] + 6));
    al((void**)&net->ws_loss, 4 * (size_t)max_batch);
    al((void**)&net->out, 4 * 6 * (size_t)max_batch);
    al((void**)&net->loss, 256);
    al((void**)&net->src_dev, sizeof(FrameSrc) * (size_t)max_batch);
    al((void**)&net->flags_dev, 4 * (size_t)max_batch);
    if (!ok) {
        for (void* p : net->allocs) hipFree(p);
        if (net->stream) hipStreamDestroy(net->stream);
        delete net;
        return set_error(RT_ERR_NOMEM, "rt_cnn_create: device allocation");
    }
    *out = net;
    return RT_OK;
}

extern "C" void rt_cnn_destroy(rt_cnn* net) {
    if (!net) return;
    hipSetDevice(net->device);
    hipStreamSynchronize(net->stream);
    net->st_src.release();
    net->st_flags.release();
    for (void* p : net->allocs) hipFree(p);
    hipStreamDestroy(net->stream);
    delete net;
}

extern "C" size_t rt_cnn_num_params(const rt_cnn* net) { return net ? net->nparams : 0; }
extern "C" void* rt_cnn_stream(rt_cnn* net) { return net ? (void*)net->stream : nullptr; }

static int pack(rt_cnn* net, hipStream_t s) {
    for (int l = 0; l < 6; l++) {
        int Co = net->cfg.channels[l + 1], Ci = net->cfg.channels[l];
        int n = Co * net->Kp[l];
        k_pack<<<(n + 255) / 256, 256, 0, s>>>(net->params + net->w_off[l], Co, Ci, net->Kp[l], net->Wp[l], net->WT[l]);
    }
    if (hipGetLastError() != hipSuccess) return set_error(RT_ERR_HIP, "pack launch");
    net->packed = true;
    return RT_OK;
}

extern "C" int rt_cnn_set_params(rt_cnn* net, const float* params) {
    if (!net || !params) return set_error(RT_ERR_ARG, "rt_cnn_set_params: null argument");
    if (hipSetDevice(net->device) != hipSuccess) return set_error(RT_ERR_HIP, "hipSetDevice");
    size_t b = 4 * net->nparams;
    if (hipMemcpyAsync(net->params, params, b, hipMemcpyHostToDevice, net->stream) != hipSuccess ||
        hipMemsetAsync(net->adam_m, 0, b, net->stream) != hipSuccess || hipMemsetAsync(net->adam_v, 0, b, net->stream) != hipSuccess)
        return set_error(RT_ERR_HIP, "rt_cnn_set_params: copy");
    net->step = 0;
    int rc = pack(net, net->stream);
    if (rc) return rc;
    if (hipStreamSynchronize(net->stream) != hipSuccess) return set_error(RT_ERR_HIP, "rt_cnn_set_params: sync");
    return RT_OK;
}

extern "C" int rt_cnn_get_params(rt_cnn* net, float* params) {
    if (!net || !params) return set_error(RT_ERR_ARG, "rt_cnn_get_params: null argument");
    if (hipSetDevice(net->device) != hipSuccess) return set_error(RT_ERR_HIP, "hipSetDevice");
    if (hipMemcpyAsync(params, net->params, 4 * net->nparams, hipMemcpyDeviceToHost, net->stream) != hipSuccess ||
        hipStreamSynchronize(net->stream) != hipSuccess)
        return set_error(RT_ERR_HIP, "rt_cnn_get_params: copy");
    return RT_OK;
}

// One forward layer on the 16x16x32 MFMA kernel when its shape fits: Cin = 4
// or a power of two >= 8, Cout a multiple of 32 (Cin = 4) or 64. The wave tile
// is the tallest of 64, 32, 16 rows that still gives 2048 waves. Returns false
// when the layer needs the general kernels.
static bool forward_mfma(const ConvGeom& g, const bf16* X, const bf16* Wp, const float* bias, bf16* Y, hipStream_t s) {
    const int M = g.B * g.Ho * g.Wo;
    const bool pow2 = g.Ci >= 8 && (g.Ci & (g.Ci - 1)) == 0;
    if (g.Ci == 4 && g.Co % 32 == 0) {
        k_conv_fwd_mfma<4, 2, 4, 1, true><<<dim3((M + 255) / 256, g.Co / 32), 256, 0, s>>>(g, X, Wp, bias, Y);
        return true;
    }
    if (!pow2 || g.Co % 64) return false;
    if (g.Ci % 32 == 0) {
        // LDS-staged tiles: the largest workgroup tile that still gives 512 workgroups
        auto blocks = [&](int bm, int bn) { return (long)((M + bm - 1) / bm) * (g.Co / bn); };
        if (g.Co == 64) {
            if (blocks(256, 64) >= 512) k_conv_fwd_lds32<256, 64, 4><<<dim3((M + 255) / 256, 1), 256, 0, s>>>(g, X, Wp, bias, Y);
            else k_conv_fwd_lds32<128, 64, 4><<<dim3((M + 127) / 128, 1), 256, 0, s>>>(g, X, Wp, bias, Y);
            return true;
        }
        if (g.Co % 128 == 0) {
            const unsigned gy = g.Co / 128;
            if (blocks(128, 128) >= 512) k_conv_fwd_lds32<128, 128, 2><<<dim3((M + 127) / 128, gy), 256, 0, s>>>(g, X, Wp, bias, Y);
            else if (blocks(64, 128) >= 512) k_conv_fwd_lds32<64, 128, 2><<<dim3((M + 63) / 64, gy), 256, 0, s>>>(g, X, Wp, bias, Y);
            else if (g.Ci % 64 == 0)  // few rows and a long k: stages of 64 halve its barriers and load round trips
                k_conv_fwd_lds32<32, 128, 2, 64><<<dim3((M + 31) / 32, gy), 256, 0, s>>>(g, X, Wp, bias, Y);
            else k_conv_fwd_lds32<32, 128, 2><<<dim3((M + 31) / 32, gy), 256, 0, s>>>(g, X, Wp, bias, Y);
            return true;
        }
    }
    auto waves = [&](int wm) { return (long)((M + wm - 1) / wm) * (g.Co / 64); };
    if (g.Co == 64) {  // one column tile: the four waves stacked over rows
        if (waves(64) >= 2048) k_conv_fwd_mfma<4, 4, 4, 1, false><<<dim3((M + 255) / 256, 1), 256, 0, s>>>(g, X, Wp, bias, Y);
        else if (waves(32) >= 2048) k_conv_fwd_mfma<2, 4, 4, 1, false><<<dim3((M + 127) / 128, 1), 256, 0, s>>>(g, X, Wp, bias, Y);
        else k_conv_fwd_mfma<1, 4, 4, 1, false><<<dim3((M + 63) / 64, 1), 256, 0, s>>>(g, X, Wp, bias, Y);
        return true;
    }
    if (g.Co % 128) return false;
    const dim3 gy(0, g.Co / 128);
    if (waves(64) >= 2048) k_conv_fwd_mfma<4, 4, 2, 2, false><<<dim3((M + 127) / 128, gy.y), 256, 0, s>>>(g, X, Wp, bias, Y);
    else if (waves(32) >= 2048) k_conv_fwd_mfma<2, 4, 2, 2, false><<<dim3((M + 63) / 64, gy.y), 256, 0, s>>>(g, X, Wp, bias, Y);
    else k_conv_fwd_mfma<1, 4, 2, 2, false><<<dim3((M + 31) / 32, gy.y), 256, 0, s>>>(g, X, Wp, bias, Y);
    return true;
}

// The first two layers in one launch when their shapes are the fused
// kernel's: 4 -> 32 -> 64 channels, the first layer's map within its LDS, the
// second layer's rows within 3 row tiles per wave.
static bool forward_fused12(rt_cnn* net, int n, hipStream_t s) {
    const ConvGeom g1 = geom(net, 0, n), g2 = geom(net, 1, n);
    if (g1.Ci != 4 || g1.Co != F12_C1 || g1.Kp != 64 || g2.Ci != F12_C1 || g2.Co != 64 || g2.Kp != 9 * F12_C1 ||
        g2.Hi != g1.Ho || g2.Wi != g1.Wo || g1.Ho * g1.Wo * F12_C1 > F12_LDS ||
        g2.Ho * g2.Wo * (64 + 8) > F12_LDS || (g2.Ho * g2.Wo + 15) / 16 > 3 * 8)
        return false;
    k_conv12_fused<64><<<n, 512, 0, s>>>(g1, g2, net->act[0], net->Wp[0], net->params + net->b_off[0], net->Wp[1],
                                         net->params + net->b_off[1], net->act[2]);
    return true;
}

// keep: every layer's activation stays in net->act (training's backward pass
// reads them); otherwise the first two layers may run fused and the first
// layer's activation is not stored.
static int forward(rt_cnn* net, int n, const bf16* x, hipStream_t s, bool keep = false) {
    if (!net->packed) return set_error(RT_ERR_STATE, "rt_cnn: parameters not set");
    net->act[0] = const_cast<bf16*>(x);
    const bool fused = !keep && forward_fused12(net, n, s);
    for (int l = fused ? 2 : 0; l < 6; l++) {
        ConvGeom g = geom(net, l, n);
        int M = n * g.Ho * g.Wo;
        if (forward_mfma(g, net->act[l], net->Wp[l], net->params + net->b_off[l], net->act[l + 1], s)) continue;
        if (g.Ci % FT_K == 0 && g.Co % FT_N == 0) {
            dim3 grid((M + FT_M - 1) / FT_M, g.Co / FT_N);
            k_conv_fwd_lds<<<grid, 256, 0, s>>>(g, net->act[l], net->Wp[l], net->params + net->b_off[l],
                                                net->act[l + 1]);
        } else {
            dim3 grid((M + 63) / 64, (g.Co + 63) / 64);
            k_conv_fwd<<<grid, 256, 0, s>>>(g, net->act[l], net->Wp[l], net->params + net->b_off[l], net->act[l + 1]);
        }
    }
    return hipGetLastError() == hipSuccess ? RT_OK : set_error(RT_ERR_HIP, "rt_cnn forward launch");
}

// forward + backward into net->grads; loss into net->loss[0] (device)
static int loss_grad(rt_cnn* net, int n, const bf16* x, const float* y, const float* wn, hipStream_t s) {
    int rc = forward(net, n, x, s, true);
    if (rc) return rc;
    k_weight_norm<<<1, 64, 0, s>>>(wn, n, net->loss + 1);
    const int C6 = net->cfg.channels[6], P6 = net->H[6] * net->W[6];
    bf16* dz = net->dz_a;
    bf16* dz_other = net->dz_b;
    k_head<<<n, 256, 2 * C6 * 4, s>>>(P6, C6, net->act[6], net->params + net->fc_w, net->params + net->fc_b, y, wn,
                                      net->loss + 1, net->out, dz, net->ws_fc, net->ws_loss);
    int width = 6 * C6 + 6;
    k_reduce_rows<<<(width + 255) / 256, 256, 0, s>>>(net->ws_fc, n, width, net->grads + net->fc_w);
    k_reduce_rows<<<1, 256, 0, s>>>(net->ws_loss, n, 1, net->loss);
    for (int l = 5; l >= 0; l--) {
        ConvGeom g = geom(net, l, n);
        int M = n * g.Ho * g.Wo;
        int S, chunk;
        wsplit(net, l, n, &S, &chunk);
        dim3 gw((g.Co + 31) / 32, g.Kp / 32, S);
        k_conv_bwd_w<<<gw, 64, 0, s>>>(g, dz, net->act[l], net->ws, chunk);
        int nw = g.Co * g.K;
        k_reduce_w<<<(nw + 255) / 256, 256, 0, s>>>(net->ws, S, g.Co, g.Kp, g.K, net->grads + net->w_off[l]);
        int Sb = std::min(1024, (M + 255) / 256);
        int cb = (M + Sb - 1) / Sb;
        Sb = (M + cb - 1) / cb;
        k_colsum<<<dim3((g.Co + 255) / 256, Sb), 256, 0, s>>>(dz, M, g.Co, cb, net->ws_b);
        k_reduce_rows<<<(g.Co + 255) / 256, 256, 0, s>>>(net->ws_b, Sb, g.Co, net->grads + net->b_off[l]);
        if (l > 0) {
            int Mi = n * g.Hi * g.Wi;
            dim3 gd((Mi + 63) / 64, (g.Ci + 63) / 64);
            k_conv_bwd_data<<<gd, 256, 0, s>>>(g, dz, net->WT[l], net->act[l], dz_other);
            std::swap(dz, dz_other);
        }
    }
    return hipGetLastError() == hipSuccess ? RT_OK : set_error(RT_ERR_HIP, "rt_cnn backward launch");
}

static int check_batch(rt_cnn* net, int n, const void* x) {
    if (!net || !x || n < 1 || n > net->max_batch) return set_error(RT_ERR_ARG, "rt_cnn: bad batch");
    if (hipSetDevice(net->device) != hipSuccess) return set_error(RT_ERR_HIP, "hipSetDevice");
    return RT_OK;
}

extern "C" int rt_cnn_forward(rt_cnn* net, int n, const uint16_t* x_dev, float* out_dev) {
    int rc = check_batch(net, n, x_dev);
    if (rc) return rc;
    if (!out_dev) return set_error(RT_ERR_ARG, "rt_cnn_forward: null output");
    if ((rc = forward(net, n, x_dev, net->stream))) return rc;
    const int C6 = net->cfg.channels[6], P6 = net->H[6] * net->W[6];
    k_head<<<n, 256, 2 * C6 * 4, net->stream>>>(P6, C6, net->act[6], net->params + net->fc_w, net->params + net->fc_b,
                                                nullptr, nullptr, nullptr, out_dev, nullptr, nullptr, nullptr);
    if (hipGetLastError() != hipSuccess) return set_error(RT_ERR_HIP, "rt_cnn head launch");
    return RT_OK;
}

extern "C" int rt_cnn_loss_grad(rt_cnn* net, int n, const uint16_t* x_dev, const float* y_dev, const float* w_dev,
                                float* loss, float* out_host, float* grad_host) {
    int rc = check_batch(net, n, x_dev);
    if (rc) return rc;
    if (!y_dev || !w_dev) return set_error(RT_ERR_ARG, "rt_cnn_loss_grad: null targets");
    if ((rc = loss_grad(net, n, x_dev, y_dev, w_dev, net->stream))) return rc;
    bool ok = true;
    if (loss) ok = ok && hipMemcpyAsync(loss, net->loss, 4, hipMemcpyDeviceToHost, net->stream) == hipSuccess;
    if (out_host) ok = ok && hipMemcpyAsync(out_host, net->out, 24 * (size_t)n, hipMemcpyDeviceToHost, net->stream) == hipSuccess;
    if (grad_host) ok = ok && hipMemcpyAsync(grad_host, net->grads, 4 * net->nparams, hipMemcpyDeviceToHost, net->stream) == hipSuccess;
    ok = ok && hipStreamSynchronize(net->stream) == hipSuccess;
    return ok ? RT_OK : set_error(RT_ERR_HIP, "rt_cnn_loss_grad: copy");
}

extern "C" int rt_cnn_train_step(rt_cnn* net, int n, const uint16_t* x_dev, const float* y_dev, const float* w_dev,
                                 float lr, float* loss) {
    int rc = check_batch(net, n, x_dev);
    if (rc) return rc;
    if (!y_dev || !w_dev || !(lr >= 0)) return set_error(RT_ERR_ARG, "rt_cnn_train_step: bad argument");
    if ((rc = loss_grad(net, n, x_dev, y_dev, w_dev, net->stream))) return rc;
    net->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    float c1 = (float)(1.0 - pow(0.9, net->step)), c2 = (float)(1.0 - pow(0.999, net->step));
    size_t P = net->nparams;
    k_adam<<<(unsigned)((P + 255) / 256), 256, 0, net->stream>>>(net->params, net->grads, net->adam_m, net->adam_v, P, lr, b1,
                                                                  b2, c1, c2, 1e-8f);
    if (hipGetLastError() != hipSuccess) return set_error(RT_ERR_HIP, "rt_cnn adam launch");
    if ((rc = pack(net, net->stream))) return rc;
    if (loss) {
        if (hipMemcpyAsync(loss, net->loss, 4, hipMemcpyDeviceToHost, net->stream) != hipSuccess ||
            hipStreamSynchronize(net->stream) != hipSuccess)
            return set_error(RT_ERR_HIP, "rt_cnn_train_step: loss copy");
    }
    return RT_OK;
}

extern "C" int rt_cnn_augment(rt_cnn* net, int n, const uint16_t* x_dev, const float* y_dev, const float* y_inv_dev,
                              const int* flags, uint16_t* x_out, float* y_out) {
    int rc = check_batch(net, n, x_dev);
    if (rc) return rc;
    if (!y_dev || !y_inv_dev || !flags || !x_out || !y_out) return set_error(RT_ERR_ARG, "rt_cnn_augment: null argument");
    if (net->cfg.channels[0] != 4) return set_error(RT_ERR_UNSUPPORTED, "rt_cnn_augment: 4-channel inputs only");
    if (!net->st_flags.upload(net->flags_dev, flags, 4 * (size_t)n, net->stream))
        return set_error(RT_ERR_HIP, "rt_cnn_augment: flags copy");
    int cells = net->cfg.in_width * net->cfg.in_height;
    k_augment<<<dim3((cells + 255) / 256, n), 256, 0, net->stream>>>(x_dev, y_dev, y_inv_dev, net->flags_dev, net->cfg.in_width,
                                                                     net->cfg.in_height, x_out, y_out);
    return hipGetLastError() == hipSuccess ? RT_OK : set_error(RT_ERR_HIP, "rt_cnn_augment launch");
}

static int make_inputs(rt_cnn* net, int n, const FrameSrc* src, bf16* x, hipStream_t s) {
    for (int i = 0; i < n; i++) {
        const FrameSrc& f = src[i];
        if (!f.g8[0] || !f.g8[1] || !f.d16[0] || !f.d16[1]) return set_error(RT_ERR_ARG, "rt_cnn: null frame");
        if (f.W < net->cfg.in_width || f.H < net->cfg.in_height || f.W % net->cfg.in_width || f.H % net->cfg.in_height)
            return set_error(RT_ERR_ARG, "rt_cnn: frame size must be a multiple of the network input size");
    }
    if (!net->st_src.upload(net->src_dev, src, sizeof(FrameSrc) * (size_t)n, s))
        return set_error(RT_ERR_HIP, "rt_cnn: frame table copy");
    int cells = net->cfg.in_width * net->cfg.in_height;
    k_inputs<<<dim3((cells + 255) / 256, n), 256, 0, s>>>(net->src_dev, net->cfg.in_width, net->cfg.in_height, x);
    return hipGetLastError() == hipSuccess ? RT_OK : set_error(RT_ERR_HIP, "rt_cnn inputs launch");
}

extern "C" int rt_cnn_make_inputs(rt_cnn* net, int n, const rt_pair* pairs, uint16_t* x_dev) {
    int rc = check_batch(net, n, x_dev);
    if (rc) return rc;
    if (!pairs) return set_error(RT_ERR_ARG, "rt_cnn_make_inputs: null pairs");
    std::vector<FrameSrc> src(n);
    for (int i = 0; i < n; i++) {
        src[i].W = pairs[i].width; src[i].H = pairs[i].height;
        src[i].g8[0] = pairs[i].ref_gray; src[i].g8[1] = pairs[i].cur_gray;
        src[i].d16[0] = pairs[i].ref_depth; src[i].d16[1] = pairs[i].cur_depth;
    }
    if ((rc = make_inputs(net, n, src.data(), x_dev, net->stream))) return rc;
    return hipStreamSynchronize(net->stream) == hipSuccess ? RT_OK : set_error(RT_ERR_HIP, "rt_cnn_make_inputs: sync");
}

namespace rt {
bool cnn_cells8(const rt_cnn* net, int W, int H) {
    return net && net->cfg.in_width * 8 == W && net->cfg.in_height * 8 == H;
}

int cnn_propose(rt_cnn* net, int n, const FrameSrc* src, float* trel_dev, hipStream_t s, const bf16* x_ready) {
    if (!net || n < 1 || n > net->max_batch) return set_error(RT_ERR_ARG, "rt_batch_propose: batch larger than the network's max_batch");
    if (!net->packed) return set_error(RT_ERR_STATE, "rt_batch_propose: network parameters not set");
    if (hipSetDevice(net->device) != hipSuccess) return set_error(RT_ERR_HIP, "hipSetDevice");
    const bf16* x = x_ready;
    int rc = 0;
    if (!x) {
        if ((rc = make_inputs(net, n, src, net->xin, s))) return rc;
        x = net->xin;
    }
    if ((rc = forward(net, n, x, s))) return rc;
    const int C6 = net->cfg.channels[6], P6 = net->H[6] * net->W[6];
    k_head<<<n, 256, 2 * C6 * 4, s>>>(P6, C6, net->act[6], net->params + net->fc_w, net->params + net->fc_b, nullptr,
                                      nullptr, nullptr, net->out, nullptr, nullptr, nullptr);
    k_outputs_to_trel<<<(n + 255) / 256, 256, 0, s>>>(net->out, n, net->cfg.rot_scale_deg * (float)(M_PI / 180.0),
                                                      net->cfg.trans_scale_mm, trel_dev);
    return hipGetLastError() == hipSuccess ? RT_OK : set_error(RT_ERR_HIP, "rt_batch_propose launch");
}
}  // namespace rt

// ---------------------------------------------------------------- host pose conversions
static void log_rotation(const double* T, double w[3]) {
    // rotation vector of the 3x3 block of a row-major 4x4
    double R[9] = {T[0], T[1], T[2], T[4], T[5], T[6], T[8], T[9], T[10]};
    double sx = R[7] - R[5], sy = R[2] - R[6], sz = R[3] - R[1];
    double s = 0.5 * sqrt(sx * sx + sy * sy + sz * sz);
    double c = 0.5 * (R[0] + R[4] + R[8] - 1.0);
    double th = atan2(s, c);
    if (th < 1e-9) { w[0] = 0.5 * sx; w[1] = 0.5 * sy; w[2] = 0.5 * sz; return; }
    if (M_PI - th < 1e-6) {
        // near pi: axis from the diagonal
        double ax[3] = {sqrt(std::max(0.0, (R[0] + 1) / 2)), sqrt(std::max(0.0, (R[4] + 1) / 2)), sqrt(std::max(0.0, (R[8] + 1) / 2))};
        if (R[1] < 0) ax[1] = -ax[1];
        if (R[2] < 0) ax[2] = -ax[2];
        for (int k = 0; k < 3; k++) w[k] = th * ax[k];
        return;
    }
    double f = th / (2.0 * s);
    w[0] = f * sx; w[1] = f * sy; w[2] = f * sz;
}

extern "C" int rt_cnn_pose_to_target(const rt_cnn_config* c, int n, const double* Trel, float* y) {
    if (!c || !Trel || !y || n < 0) return set_error(RT_ERR_ARG, "rt_cnn_pose_to_target: bad argument");
    double rs = c->rot_scale_deg * M_PI / 180.0;
    for (int i = 0; i < n; i++) {
        const double* T = Trel + 16 * (size_t)i;
        double w[3];
        log_rotation(T, w);
        for (int k = 0; k < 3; k++) {
            y[6 * i + k] = (float)(w[k] / rs);
            y[6 * i + 3 + k] = (float)(T[4 * k + 3] / c->trans_scale_mm);
        }
    }
    return RT_OK;
}

extern "C" int rt_cnn_target_to_pose(const rt_cnn_config* c, int n, const float* y, double* Trel) {
    if (!c || !Trel || !y || n < 0) return set_error(RT_ERR_ARG, "rt_cnn_target_to_pose: bad argument");
    double rs = c->rot_scale_deg * M_PI / 180.0;
    for (int i = 0; i < n; i++) {
        double w[3], t[3];
        for (int k = 0; k < 3; k++) { w[k] = y[6 * i + k] * rs; t[k] = (double)y[6 * i + 3 + k] * c->trans_scale_mm; }
        double th = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
        double A = th > 1e-12 ? sin(th) / th : 1.0, B = th > 1e-12 ? (1.0 - cos(th)) / (th * th) : 0.5;
        double K[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
        double* T = Trel + 16 * (size_t)i;
        for (int r = 0; r < 3; r++) {
            for (int q = 0; q < 3; q++) {
                double k2 = 0.0;
                for (int m = 0; m < 3; m++) k2 += K[3 * r + m] * K[3 * m + q];
                T[4 * r + q] = (r == q ? 1.0 : 0.0) + A * K[3 * r + q] + B * k2;
            }
            T[4 * r + 3] = t[r];
        }
        T[12] = 0; T[13] = 0; T[14] = 0; T[15] = 1;
    }
    return RT_OK;
}
