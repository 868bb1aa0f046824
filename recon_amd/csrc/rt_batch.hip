// Batched tracking of many independent frame pairs of mixed sizes.
//
// Keyframe preparation runs as batched versions of the single-pair kernels
// (grid z or y = pair), sharing their device bodies (rt_prep.hpp), so every
// per-pixel and per-point value is the same as on the single-pair path. The
// Gauss-Newton loop runs in ONE persistent launch: each workgroup takes pairs
// from a device work queue (ordered by decreasing cost) and runs the whole
// coarse-to-fine photometric loop of a pair with workgroup barriers only:
// residuals stay in registers between the residual and the system pass, the MAD
// histogram lives in LDS, the 27 normal-equation terms are reduced in FP32 per
// wave and combined across waves in FP64, and one lane runs the FP64 solve and
// SE(3) update. Workgroup size and points per thread are template parameters
// picked per batch by a timing autotuner; the pipeline is captured once into a
// hipGraph and replayed.
#include <hip/hip_runtime.h>
#include <type_traits>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <vector>
#include "../../include/rt_tracker.h"
#include "cnn.hpp"
#include "rt_kernels.hpp"
#include "rt_prep.hpp"
#include "rt_host.hpp"
#include "rt_stage.hpp"

namespace rt {

struct PairDesc {
    int W, H, L, N;
    int iters[4];
    const float* calib;
    int simple;             // calib_simple(calib), evaluated at preparation
    float ifx, ify;         // RN(1 / fx), RN(1 / fy) of the depth camera (IEEE divisions on the host)
    bf16* feat;             // the pose network's inputs of the pair (cells of 8 x 8 pixels, 4 channels), or null
    uint16_t* zblk;         // the current depth map in 8 x 8-pixel blocks of 64 values (one 128-byte line each)
    const uint8_t* g8[2];
    const uint16_t* d16[2];
    float* gray[2][4];
    float* z[2];
    uint8_t* bins;          // per-pixel gradient-magnitude bin
    unsigned* ghist;
    int* sel;
    float4* pts;
    float4* jac[4];
    const float* trel_in;   // 16
    float* trel_out;        // 16
    float* rms;             // 1
    int* status;            // 1
    // geometric and bi-objective modes: keyframe cloud (x, y, z, mask) and
    // normals, current cloud (x, y, z, valid) per level; the compacted ICP
    // points and squared residuals of the level being tracked; per level the
    // ICP point count and the three Student-t variance sums
    float4* rpos[4];
    float4* rnrm[4];
    float4* cpos[4];
    int* isel;
    float* ir2;
    int* icnt;              // 4
    float* ivar;            // 4 x 3
};

// ------------------------------------------------------------------ batched preparation
// Level 0 of both frames stays in its input form (8-bit gray, 16-bit depth):
// every consumer reads it through the exact conversions (U8Src, U16DepthSrc),
// so no FP32 copy is written or read back.
// (float)t / 255 for every 8-bit value, evaluated by the compiler (IEEE
// division, round to nearest: the values the device division gives)
struct U8Table { float v[256]; };
constexpr U8Table make_u8_table() {
    U8Table l{};
    for (int t = 0; t < 256; t++) l.v[t] = (float)t / 255.0f;
    return l;
}
__constant__ U8Table c_u8_table = make_u8_table();
__device__ __forceinline__ void load_lut(float* lut) {
    for (int t = threadIdx.x; t < 256; t += blockDim.x) lut[t] = c_u8_table.v[t];
}

// pyramid levels 1..L-1 of one frame (fastDownSampleHdrKernel, downSample2Cuda.cu:35-42);
// the keyframe's (which 0) are built by kb_bins from its gray tile
// The current frame (which 1) also gets its depth map copied into 8 x 8-pixel
// blocks for the tracker, and with d.feat set the pose network's inputs over
// the block's 8 x 8 cells: four threads per cell, two columns each, column
// sums over the rows in order, then pairwise (the order of cell8_features,
// the network's input builder)
__global__ __launch_bounds__(256) void kb_pyramid(const PairDesc* __restrict__ P, int which) {
    const PairDesc& d = P[blockIdx.z];
    if ((int)blockIdx.x * 64 >= d.W || (int)blockIdx.y * 64 >= d.H) return;
    __shared__ float lut[256];
    load_lut(lut);
    __syncthreads();
    convert_pyramid_block<3>(blockIdx.x, blockIdx.y, d.g8[which], nullptr, nullptr, d.W, d.H, d.L, d.calib, nullptr,
                             d.gray[which][1], d.gray[which][2], d.gray[which][3], nullptr, lut);
    if (which != 1) return;
    bf16* const feat = d.feat;
    const int W = d.W, H = d.H;
    const int cell = threadIdx.x >> 2, part = threadIdx.x & 3;
    const int cx = blockIdx.x * 8 + (cell & 7), cy = blockIdx.y * 8 + (cell >> 3);
    // columns clamped into the image (W is a multiple of 8), rows one by one;
    // what lies outside is neither stored nor summed into a written cell
    const unsigned px = min(cx * 8 + 2 * part, W - 2);
    const RT_GLOBAL uint8_t* g8 = gmem(d.g8[1]);
    const RT_GLOBAL uint16_t* d16 = gmem(d.d16[1]);
    const bool d4 = ((uintptr_t)d.d16[1] & 3) == 0;  // two depths per 4-byte load
    unsigned gw[8], dw[8];
#pragma unroll
    for (int r = 0; r < 8; r++) {
        const unsigned o = px + __umul24(min(cy * 8 + r, H - 1), W);
        if (feat) gw[r] = at_offset((const RT_GLOBAL uint16_t*)g8, o);  // two gray bytes (o even)
        dw[r] = d4 ? at_offset((const RT_GLOBAL unsigned*)d16, 2 * o)
                   : (unsigned)d16[o] | (unsigned)d16[o + 1] << 16;
    }
    // the current depth map in 8 x 8-pixel blocks (the tracker's depth taps):
    // block (cx, cy) is 32 words, row r of it words 4r .. 4r + 3, two depths
    // each. The region's 64 blocks are assembled in LDS, then each row of 8
    // blocks (consecutive in memory) is stored as one 1 KB run, 32 bytes per
    // thread. Rows past H (when H is not a multiple of 8) hold copies of row
    // H - 1 and are never read.
    __shared__ __align__(16) unsigned zt[64 * 32];
#pragma unroll
    for (int r = 0; r < 8; r++) zt[cell * 32 + 4 * r + part] = dw[r];
    __syncthreads();
    {
        const int W8 = W >> 3, H8 = (H + 7) >> 3;
        const int cyl = threadIdx.x >> 5, w = (threadIdx.x & 31) * 8;  // cell row, first of 8 words
        const int bcx = blockIdx.x * 8 + (w >> 5), bcy = blockIdx.y * 8 + cyl;
        if (bcx < W8 && bcy < H8) {
            typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
            const u32x4* src = (const u32x4*)&zt[cyl * 256 + w];
            RT_GLOBAL u32x4* dst = (RT_GLOBAL u32x4*)((RT_GLOBAL unsigned*)gmem(d.zblk) + 32 * ((size_t)bcy * W8 + bcx) + (w & 31));
            dst[0] = src[0];
            dst[1] = src[1];
        }
    }
    if (!feat) return;
    const bool valid = cx < (W >> 3) && cy < (H >> 3);
    unsigned fg = 0, fn = 0;
    float c0 = 0.0f, c1 = 0.0f;
#pragma unroll
    for (int r = 0; r < 8; r++) {
        fg += (gw[r] & 0xffu) + (gw[r] >> 8);
        const unsigned z0 = dw[r] & 0xffffu, z1 = dw[r] >> 16;
        if (z0) { c0 += inv_depth_m(z0); fn++; }
        if (z1) { c1 += inv_depth_m(z1); fn++; }
    }
    // thread part p holds columns 2p, 2p + 1: (c0 + c1) + (c2 + c3), then + (c4 + c5) + (c6 + c7)
    float fi = c0 + c1;
#pragma unroll
    for (int o = 1; o < 4; o <<= 1) {
        fi += __shfl_xor(fi, o, 64);
        fg += __shfl_xor(fg, o, 64);
        fn += __shfl_xor(fn, o, 64);
    }
    if (part == 0 && valid) cell_features(fg, fi, fn, 64, feat + 4 * ((size_t)cy * (W >> 3) + cx) + 2);
}
// clears each pair's gradient histogram before kb_bins counts into it
__global__ __launch_bounds__(256) void kb_clear_hist(const PairDesc* __restrict__ P) {
    P[blockIdx.x].ghist[threadIdx.x] = 0u;
}

// xyz2CloudKernel's gradient magnitude (d2z.cu:235-331) as an 8-bit bin per
// pixel, and the 256-bin histogram (cuda_hist.cu:100-152). Tiles of 64 x 16
// (one wave per 4 rows): the tile's histogram is counted in LDS (one LDS
// atomic per pixel) and merged into the pair's with one global atomic per
// non-empty bin.
constexpr int BINS_TILE_H = 32;
__global__ __launch_bounds__(256) void kb_bins(const PairDesc* __restrict__ P) {
    const PairDesc& d = P[blockIdx.z];
    if ((int)blockIdx.x * 64 >= d.W || (int)blockIdx.y * BINS_TILE_H >= d.H) return;
    // tile plus a halo of 2 rows and 4 columns (whole 4-byte words: the
    // width is a multiple of 8, so a word lies entirely inside or outside a row)
    constexpr int TW = 64 + 8, TH = BINS_TILE_H + 4, TWW = TW / 4, NW = TWW * TH;
    static_assert(TW == 4 * TWW && BINS_TILE_H % 8 == 0, "tile of whole words; three pyramid levels");
    constexpr int ROWS = BINS_TILE_H / 4;  // pixels per thread, one per row
    __shared__ float lut[256];
    // four histogram copies, by lane: lanes with equal bins mostly update
    // different addresses (neighbouring pixels tend to share a bin)
    constexpr int HC = 4;
    __shared__ unsigned h[256 * HC];
    __shared__ float gt[TW * TH];
    const int W = d.W, H = d.H;
    // the calibration, read before the kernel's first store: wave-uniform
    // loads of memory nothing here writes, so they become scalar loads into
    // scalar registers
    float cl[96];
    const RT_GLOBAL float* cg = gmem(d.calib);
#pragma unroll
    for (int k = 0; k < 96; k++) cl[k] = cg[k];
    const int lane = lane_id(), wave = threadIdx.x >> 6;
    const int x = blockIdx.x * 64 + lane;
    const int y0 = blockIdx.y * BINS_TILE_H + wave * ROWS;
    const RT_GLOBAL uint16_t* d16 = gmem(d.d16[0]);
    const RT_GLOBAL uint8_t* g8 = gmem(d.g8[0]);
    const int tx0 = (int)blockIdx.x * 64 - 4, ty0 = (int)blockIdx.y * BINS_TILE_H - 2;
    // Every global load of the block is issued before the first wait: the
    // loads are unconditional, at coordinates clamped into the image, and
    // what lies outside is discarded afterwards.
    const int xc = min(x, W - 1);
    unsigned zr[ROWS];
#pragma unroll
    for (int k = 0; k < ROWS; k++) zr[k] = at_offset(d16, 2 * (xc + __umul24(min(y0 + k, H - 1), W)));
    // With d.feat set (W and H multiples of 8), the pose network's keyframe
    // inputs too: each wave's 8 rows x 64 columns are 8 cells of 8 x 8 pixels,
    // the lane's column sums taken over its rows in order (cell_features)
    static_assert(ROWS == 8, "one cell row per wave");
    bf16* const feat = d.feat;
    unsigned gr[ROWS];
    if (feat) {
#pragma unroll
        for (int k = 0; k < ROWS; k++) gr[k] = at_offset(g8, xc + __umul24(min(y0 + k, H - 1), W));
    }
    const bool aligned = ((uintptr_t)d.g8[0] & 3) == 0;
    constexpr int NR = (NW + 255) / 256;
    unsigned wv[NR];
    bool win[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) {
        const int e = min((int)threadIdx.x + r * 256, NW - 1);
        const int gx = tx0 + 4 * (e % TWW), gy = ty0 + e / TWW;
        win[r] = gx >= 0 && gx < W && gy >= 0 && gy < H;
        const unsigned o = (unsigned)min(max(gx, 0), W - 4) + __umul24(min(max(gy, 0), H - 1), W);
        if (aligned) wv[r] = at_offset((const RT_GLOBAL unsigned*)g8, o);
        else wv[r] = g8[o] | (unsigned)g8[o + 1] << 8 | (unsigned)g8[o + 2] << 16 | (unsigned)g8[o + 3] << 24;
    }
#pragma unroll
    for (int r = 0; r < NR; r++) wv[r] = win[r] ? wv[r] : 0u;
    load_lut(lut);
#pragma unroll
    for (int k = 0; k < HC; k++) h[threadIdx.x + 256 * k] = 0u;
    __syncthreads();
    // words outside the image are 0, and lut[0] = 0: those tile entries are
    // never read (the bounds test keeps every tap 9 pixels inside the image)
#pragma unroll
    for (int r = 0; r < NR; r++) {
        const int e = threadIdx.x + r * 256;
        if (e < NW) {
            float* q = gt + 4 * e;  // row e / TWW, column 4 (e % TWW): TW = 4 TWW
#pragma unroll
            for (int j = 0; j < 4; j++) q[j] = lut[(wv[r] >> (8 * j)) & 0xffu];
        }
    }
    __syncthreads();
    // the keyframe's pyramid levels 1..L-1 from the tile's core (64 x TH0
    // level-0 pixels: 32 x TH0/2, 16 x TH0/4, 8 x TH0/8), in
    // convert_pyramid_block's order ((top left + top right) + bottom left) +
    // bottom right, then / 4
    if (d.L > 1) {
        constexpr int TH0 = BINS_TILE_H;
        __shared__ float l1s[TH0 / 2][32];
        __shared__ float l2s[TH0 / 4][16];
        const int W1 = W >> 1, H1 = H >> 1, W2 = W >> 2, H2 = H >> 2;
#pragma unroll
        for (int e = threadIdx.x; e < 32 * (TH0 / 2); e += 256) {
            const int i = e & 31, j = e >> 5;
            const float* q = gt + (2 + 2 * j) * TW + 4 + 2 * i;
            const float v = (q[0] + q[1] + q[TW] + q[TW + 1]) / 4.0f;
            l1s[j][i] = v;
            const int X1 = blockIdx.x * 32 + i, Y1 = blockIdx.y * (TH0 / 2) + j;
            if (X1 < W1 && Y1 < H1) store_at_offset(gmem(d.gray[0][1]), 4 * ((unsigned)X1 + __umul24(Y1, W1)), v);
        }
        if (d.L > 2) {
            __syncthreads();
            if (threadIdx.x < 16 * (TH0 / 4)) {
                const int i = threadIdx.x & 15, j = threadIdx.x >> 4;
                const float v = (l1s[2 * j][2 * i] + l1s[2 * j][2 * i + 1] + l1s[2 * j + 1][2 * i] + l1s[2 * j + 1][2 * i + 1]) / 4.0f;
                l2s[j][i] = v;
                const int X2 = blockIdx.x * 16 + i, Y2 = blockIdx.y * (TH0 / 4) + j;
                if (X2 < W2 && Y2 < H2) store_at_offset(gmem(d.gray[0][2]), 4 * ((unsigned)X2 + __umul24(Y2, W2)), v);
            }
            if (d.L > 3) {
                __syncthreads();
                if (threadIdx.x < 8 * (TH0 / 8)) {
                    const int i = threadIdx.x & 7, j = threadIdx.x >> 3;
                    const float v = (l2s[2 * j][2 * i] + l2s[2 * j][2 * i + 1] + l2s[2 * j + 1][2 * i] + l2s[2 * j + 1][2 * i + 1]) / 4.0f;
                    const int X3 = blockIdx.x * 8 + i, Y3 = blockIdx.y * (TH0 / 8) + j;
                    if (X3 < (W >> 3) && Y3 < (H >> 3)) store_at_offset(gmem(d.gray[0][3]), 4 * ((unsigned)X3 + __umul24(Y3, W >> 3)), v);
                }
            }
        }
    }
    const LdsTile tile{gt, tx0, ty0, TW, TH};
    const U16DepthSrc zs{d.d16[0]};
    const U8Src gs{d.g8[0], lut};
    const GradConst gk{d.ifx, d.ify, (float)(W - 10), (float)(H - 10)};
    RT_GLOBAL uint8_t* bins = gmem(d.bins);
    auto pixels = [&](auto simple) {
#pragma unroll
        for (int k = 0; k < ROWS; k++) {
            const int y = y0 + k;
            if (x < W && y < H) {
                const float g = grad_mag_tiled<decltype(simple)::value>(-(float)zr[k], zs, gs, tile, W, H, x, y, cl, gk);
                unsigned b = grad_bin(g);
                store_at_offset(bins, (unsigned)x + __umul24(y, W), (uint8_t)b);
                atomicAdd(&h[b * HC + (lane & (HC - 1))], 1u);
            }
        }
    };
    if (d.simple) pixels(std::true_type{});
    else pixels(std::false_type{});
    if (feat) {
        unsigned fg = 0, fn = 0;
        float fi = 0.0f;
#pragma unroll
        for (int k = 0; k < ROWS; k++) {
            fg += gr[k];
            if (zr[k]) { fi += inv_depth_m(zr[k]); fn++; }
        }
        // columns 8c .. 8c+7 of a cell in lanes 8c .. 8c+7: pairwise over lanes
#pragma unroll
        for (int o = 1; o < 8; o <<= 1) {
            fi += __shfl_xor(fi, o, 64);
            fg += __shfl_xor(fg, o, 64);
            fn += __shfl_xor(fn, o, 64);
        }
        if ((lane & 7) == 0 && x < W && y0 < H)
            cell_features(fg, fi, fn, 64, feat + 4 * ((size_t)(y0 >> 3) * (W >> 3) + (x >> 3)));
    }
    __syncthreads();
    unsigned v = 0u;
#pragma unroll
    for (int k = 0; k < HC; k++) v += h[threadIdx.x * HC + k];
    if (v) atomicAdd(&d.ghist[threadIdx.x], v);
}

// Eight bins of one lane (pixels base .. base+7 of the segment [b0, end)) as
// two 32-bit words. The loads are unconditional, from addresses clamped into
// the segment, so that all of a wave's loads can be in flight together;
// bytes beyond the segment are masked off by valid_bytes. vec: the segment
// length is a multiple of 8, so one aligned 8-byte load covers them.
struct Bins8 { unsigned lo, hi; };
template <bool VEC>
__device__ __forceinline__ Bins8 load_bins8(const RT_GLOBAL uint8_t* bins, int base, int end) {
    Bins8 r;
    if constexpr (VEC) {
        typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
        u32x2 w = *(const RT_GLOBAL u32x2*)(bins + min(base, end - 8));
        r.lo = w.x; r.hi = w.y;
    } else {
        unsigned v[8];
#pragma unroll
        for (int j = 0; j < 8; j++) v[j] = bins[min(base + j, end - 1)];
        r.lo = v[0] | v[1] << 8 | v[2] << 16 | v[3] << 24;
        r.hi = v[4] | v[5] << 8 | v[6] << 16 | v[7] << 24;
    }
    return r;
}
// the lane's 8 bins at offset o of segments s0, s0 + SEL_WAVES, ...; the
// uniform branch sits outside, so each path issues all its loads back to back
// Workgroup size, segments in flight per wave and waves per SIMD of the
// selection. 512 threads with two segments in flight fit in 64 VGPRs, so four
// workgroups share a CU: measured 0.5 ms faster per 4096 pairs at 640x480
// than one 1024-thread workgroup with four segments in flight (106 VGPRs).
#ifndef RT_SEL_BS
#define RT_SEL_BS 512
#define RT_SEL_U 2
#define RT_SEL_WPE 8
#endif
// SEL_BS >= 512: one thread per segment in the scans over the 512 segments
constexpr int SEL_BS = RT_SEL_BS, SEL_WAVES = SEL_BS / 64, SEL_U = RT_SEL_U;
static_assert(SEL_BS >= 512 && SEL_BS % 64 == 0, "kb_select needs a thread per segment");
static_assert(512 % (SEL_WAVES * SEL_U) == 0, "the waves' segment groups tile the 512 segments");
__device__ __forceinline__ void load_segments(Bins8* w, const RT_GLOBAL uint8_t* bins, int s0, int Pn, int o, bool vec) {
    if (vec) {
#pragma unroll
        for (int u = 0; u < SEL_U; u++) {
            const int b0 = (s0 + u * SEL_WAVES) * Pn;
            w[u] = load_bins8<true>(bins, b0 + o, b0 + Pn);
        }
    } else {
#pragma unroll
        for (int u = 0; u < SEL_U; u++) {
            const int b0 = (s0 + u * SEL_WAVES) * Pn;
            w[u] = load_bins8<false>(bins, b0 + o, b0 + Pn);
        }
    }
}
// bit 7 of each of the first nv bytes of the lane's 8 (nv clamped to 0..8)
__device__ __forceinline__ void valid_bytes(int nv, unsigned& vlo, unsigned& vhi) {
    nv = max(0, min(nv, 8));
    const int nl = min(nv, 4), nh = nv - nl;
    vlo = nl == 4 ? 0x80808080u : 0x80808080u & ((1u << (8 * nl)) - 1u);
    vhi = nh == 4 ? 0x80808080u : 0x80808080u & ((1u << (8 * nh)) - 1u);
}
// Byte-wise comparisons of a word with a threshold (SWAR; no carry or borrow
// crosses a byte), the result in bit 7 of each byte.
// bytes equal to t: T = t replicated
__device__ __forceinline__ unsigned bytes_eq(unsigned w, unsigned T) {
    const unsigned x = w ^ T;
    const unsigned y = (x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu;  // bit 7 set where the low 7 bits of x are not all 0
    return ~(y | x | 0x7F7F7F7Fu);
}
// bytes >= u: U = u replicated (0 < u < 256). The low 7 bits compare through
// (w | 0x80) - (u & 0x7f), whose bit 7 is (w & 0x7f) >= (u & 0x7f); the top
// bits decide when they differ.
__device__ __forceinline__ unsigned bytes_ge(unsigned w, unsigned U) {
    const unsigned d = (w | 0x80808080u) - (U & 0x7F7F7F7Fu);
    return ((w & ~U) | (~(w ^ U) & d)) & 0x80808080u;
}
// inclusive sum over the lanes of a wave
__device__ __forceinline__ int wave_scan_incl(int v) {
    const int lane = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    return v;
}
// inclusive maximum over the lanes of a wave
__device__ __forceinline__ int wave_scan_max(int v) {
    const int lane = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int u = __shfl_up(v, o, 64);
        if (lane >= o) v = max(v, u);
    }
    return v;
}
__device__ __forceinline__ int wave_isum(int v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Point selection of one pair in one workgroup (seekThreshold256Kernel2 +
// filterIndexKernel4, filterIndex256.cu:26-46,110-169): threshold bin and
// excess from the histogram; per 512-pixel-group segment the counts of bins
// above / equal to the threshold; the trim of the equal counts (Q1) and the
// segment start offsets; then the ordered scatter. The segment tables stay in
// LDS. Each wave takes every SEL_WAVES-th segment, SEL_U of them at a time so
// that their loads are in flight together. Per-lane counts are packed as
// above + (equal << 16): a segment holds at most 1024 pixels (checked at
// preparation), so neither half overflows.
// above- and equal-threshold flags (bit 7 of each byte) of the lane's 8 bins
struct SelFlags { unsigned alo, ahi, elo, ehi; };
__device__ __forceinline__ SelFlags sel_flags(Bins8 w, int nv, unsigned T, unsigned U, bool any_above) {
    unsigned vlo, vhi;
    valid_bytes(nv, vlo, vhi);
    SelFlags f;
    f.alo = any_above ? bytes_ge(w.lo, U) & vlo : 0u;
    f.ahi = any_above ? bytes_ge(w.hi, U) & vhi : 0u;
    f.elo = bytes_eq(w.lo, T) & vlo;
    f.ehi = bytes_eq(w.hi, T) & vhi;
    return f;
}
__global__ __launch_bounds__(SEL_BS, RT_SEL_WPE) void kb_select(const PairDesc* __restrict__ P) {
    const PairDesc& d = P[blockIdx.x];
    __shared__ int cA[512], cE[512], sA[512], sE[512], pm[512];
    __shared__ int wtot[3][8];
    __shared__ int s_first;
    int thr_i, excess;
    find_threshold(d.ghist, d.N, thr_i, excess);
    const unsigned T = (unsigned)thr_i * 0x01010101u;
    const bool any_above = thr_i < 255;                      // no bin lies above 255
    const unsigned U = (unsigned)(thr_i + 1) * 0x01010101u;  // "above" = ">= thr + 1"
    const int t = threadIdx.x, wave = t >> 6, lane = lane_id();
    const int npix = d.W * d.H, Pn = npix / 512, N = d.N;
    const bool vec = (Pn & 7) == 0;
    const RT_GLOBAL uint8_t* bins = gmem(d.bins);
    if (t == 0) s_first = 512;
    // pass 1: counts
    for (int s0 = wave; s0 < 512; s0 += SEL_WAVES * SEL_U) {
        int ae[SEL_U];
#pragma unroll
        for (int u = 0; u < SEL_U; u++) ae[u] = 0;
        for (int off = 0; off < Pn; off += 512) {
            Bins8 w[SEL_U];
            load_segments(w, bins, s0, Pn, off + 8 * lane, vec);
#pragma unroll
            for (int u = 0; u < SEL_U; u++) {
                SelFlags f = sel_flags(w[u], Pn - off - 8 * lane, T, U, any_above);
                ae[u] += __popc(f.alo) + __popc(f.ahi) + ((__popc(f.elo) + __popc(f.ehi)) << 16);
            }
        }
#pragma unroll
        for (int u = 0; u < SEL_U; u++) {
            int v = wave_isum(ae[u]);
            if (lane == 0) { cA[s0 + u * SEL_WAVES] = v & 0xffff; cE[s0 + u * SEL_WAVES] = v >> 16; }
        }
    }
    __syncthreads();
    // Trim of the equal counts (filterIndex256.cu:141-152). Sequentially, over
    // segments j = 0..510 with a positive equal count m_j and a running excess
    // T (initially the excess): if m_j >= T the segment loses T and the trim
    // ends; otherwise the segment is emptied and T becomes m_j (the
    // reference's update T += m_j - T). In parallel: T before j is the count
    // of the previous non-empty segment (or the excess), the trim ends at the
    // first non-empty j with m_j >= T, every non-empty segment before it is
    // emptied; the last segment is never trimmed.
    int m = 0, prevval = 0;
    bool nz = false;
    if (t < 512) {
        m = cE[t];
        nz = t < 511 && m > 0;
        int v = wave_scan_max(nz ? t : -1);
        if (lane == 63) wtot[0][wave] = v;
        pm[t] = v;
    }
    __syncthreads();
    if (t < 512) {
        int prev = -1;
        for (int w = 0; w < wave; w++) prev = max(prev, wtot[0][w]);
        if (t > 0) prev = max(prev, (t - 1) >> 6 == wave ? pm[t - 1] : -1);
        prevval = prev >= 0 ? cE[prev] : excess;
        if (excess > 0 && nz && m >= prevval) atomicMin(&s_first, t);
    }
    __syncthreads();
    if (t < 512 && excess > 0 && nz && t <= s_first) m = t < s_first ? 0 : m - prevval;
    // exclusive offsets: above-threshold indices first, then the kept equal ones
    int ia = 0, ie = 0;
    if (t < 512) {
        ia = wave_scan_incl(cA[t]);
        ie = wave_scan_incl(m);
        if (lane == 63) { wtot[1][wave] = ia; wtot[2][wave] = ie; }
    }
    __syncthreads();
    if (t < 512) {
        int oa = 0, oe = 0, ta = 0;
        for (int w = 0; w < 8; w++) {
            if (w < wave) { oa += wtot[1][w]; oe += wtot[2][w]; }
            ta += wtot[1][w];
        }
        sA[t] = oa + ia - cA[t];
        sE[t] = ta + oe + ie - m;
        cE[t] = m;
    }
    int total = 0;
    for (int w = 0; w < 8; w++) total += wtot[1][w] + wtot[2][w];
    __syncthreads();
    // pass 2: ordered scatter; a lane's hits are ranked by a wave scan of the
    // packed counts, then written in pixel order (bit 8j+7 is pixel j)
    RT_GLOBAL int* sel = gmem(d.sel);
    for (int s0 = wave; s0 < 512; s0 += SEL_WAVES * SEL_U) {
        int run[SEL_U];
#pragma unroll
        for (int u = 0; u < SEL_U; u++) run[u] = 0;
        for (int off = 0; off < Pn; off += 512) {
            Bins8 w[SEL_U];
            load_segments(w, bins, s0, Pn, off + 8 * lane, vec);
#pragma unroll
            for (int u = 0; u < SEL_U; u++) {
                const int s = s0 + u * SEL_WAVES;
                const int base = s * Pn + off + 8 * lane;
                const SelFlags f = sel_flags(w[u], Pn - off - 8 * lane, T, U, any_above);
                const uint64_t am = f.alo | (uint64_t)f.ahi << 32, em = f.elo | (uint64_t)f.ehi << 32;
                const int v = __popcll(am) + (__popcll(em) << 16);
                const int incl = wave_scan_incl(v);
                const int ex = run[u] + incl - v;  // ranks of this lane's first hits, packed
                int pa = ex & 0xffff, pe = ex >> 16;
                const int startA = sA[s], startE = sE[s], mE = cE[s];
                for (uint64_t b = am; b; b &= b - 1u) {
                    int pos = startA + pa;
                    if (pos < N) sel[pos] = base + (__builtin_ctzll(b) >> 3);
                    pa++;
                }
                for (uint64_t b = em; b; b &= b - 1u) {
                    if (pe < mE) { int pos = startE + pe; if (pos < N) sel[pos] = base + (__builtin_ctzll(b) >> 3); }
                    pe++;
                }
                run[u] += __shfl(incl, 63, 64);
            }
        }
    }
    // a short selection leaves zeros behind the selected indices
    for (int k = (total > 0 ? total : 0) + t; k < N; k += SEL_BS) sel[k] = 0;
}
// XCD-aware: workgroup b runs on XCD b % 8 (dispatch is round-robin), and
// all chunks of pair p get workgroups with b % 8 == p % 8, so that the taps a
// pair's points share are fetched into one L2 instead of eight. Grid:
// 8 x ceil(n / 8) x chunks workgroups.
__global__ __launch_bounds__(256) void kb_ref_points(const PairDesc* __restrict__ P, int n, int chunks) {
    const int b = blockIdx.x, k = b >> 3;
    const int pair = (b & 7) + 8 * (k / chunks), chunk = k % chunks;
    if (pair >= n) return;
    const PairDesc& d = P[pair];
    __shared__ float lut[256];
    __shared__ float cal[96];
    load_lut(lut);
    if (threadIdx.x < 96) cal[threadIdx.x] = d.calib[threadIdx.x];
    __syncthreads();
    int i = chunk * 256 + threadIdx.x;
    if (i >= d.N) return;
    float cl[96];
    uniform_calib(cal, cl);
    ref_point(i, U16DepthSrc{d.d16[0]}, U8Src{d.g8[0], lut}, d.gray[0][1], d.gray[0][2], d.gray[0][3], d.W, d.H, d.L,
              d.N, cl, 1e-3f, d.sel, d.pts, d.jac[0], d.jac[1], d.jac[2], d.jac[3]);
}
// Geometric modes: the keyframe cloud with normals and the current cloud at
// level 0 from the 16-bit depths (ShapeTracker::setReferencePoints /
// generatePyramids, shapetracker.cpp:195-211,254-258), then the masked 2x2
// pyramids level by level. The per-pixel bodies are the single-pair kernels'
// (ref_cloud_pixel, cur_cloud_pixel, cloud_down_pixel), reading the depth
// through the same exact conversion, so the clouds are the same bit for bit.
__global__ __launch_bounds__(256) void kb_clouds(const PairDesc* __restrict__ P) {
    const PairDesc& d = P[blockIdx.z];
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= d.W || y >= d.H) return;
    ref_cloud_pixel(U16DepthSrc{d.d16[0]}, d.W, d.H, x, y, d.calib, d.rpos[0], d.rnrm[0]);
    cur_cloud_pixel(U16DepthSrc{d.d16[1]}, d.W, x + y * d.W, d.calib, d.cpos[0]);
}
__global__ __launch_bounds__(256) void kb_clouds_down(const PairDesc* __restrict__ P, int level) {
    const PairDesc& d = P[blockIdx.y];
    if (level >= d.L) return;
    const int Wd = d.W >> level, Hd = d.H >> level;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= Wd * Hd) return;
    cloud_down_pixel(d.rpos[level - 1], d.rnrm[level - 1], Wd, i, d.rpos[level], d.rnrm[level]);
    cloud_down_pixel(d.cpos[level - 1], nullptr, Wd, i, d.cpos[level], nullptr);
}

__global__ void kb_reset_queue(int* q) {
    if (threadIdx.x == 0) q[0] = 0;
}

// ------------------------------------------------------------------ geometric terms in one workgroup
// ICP selection of one level (icpResidualKernel2 + selectValidIndex +
// packIndex, rgb2gray.cu:223-260,827-850): ordered compaction of the valid
// keyframe pixels into d.isel / d.ir2 (squared residual), BS pixels a step:
// a ballot per wave, the waves' counts in LDS. Ascending pixel order, as the
// single-pair compaction. Returns the count (the same in every thread).
template <int BS>
__device__ int icp_select_wg(const PairDesc& d, int layer, const float* T, const float* c) {
    constexpr int NW = BS / 64;
    __shared__ int wc[NW];
    const int t = threadIdx.x, wave = t >> 6, lane = lane_id();
    const int w = d.W >> layer, h = d.H >> layer, npix = w * h;
    const float a = 1.0f / (float)(1 << layer), b = 0.5f * (a - 1.0f);
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    RT_GLOBAL int* isel = gmem(d.isel);
    RT_GLOBAL float* ir2 = gmem(d.ir2);
    int base = 0;
    for (int o0 = 0; o0 < npix; o0 += BS) {
        const int o = o0 + t;
        float r = 0.0f;
        const bool valid = o < npix && icp_mask_pixel(d.rpos[layer], d.rnrm[layer], d.cpos[layer], w, h, o, T, c, a, b, r);
        const uint64_t bl = __ballot(valid);
        if (lane == 0) wc[wave] = __popcll(bl);
        __syncthreads();
        int off = 0, tot = 0;
#pragma unroll
        for (int q = 0; q < NW; q++) {
            const int cq = wc[q];
            off += q < wave ? cq : 0;
            tot += cq;
        }
        if (valid) {
            const int k = base + off + __popcll(bl & lt);
            isel[k] = o;
            ir2[k] = r * r;
        }
        base += tot;
        __syncthreads();
    }
    return base;
}
// generateStudentTWeights (rgb2gray.cu:470-568, 854-877) in one workgroup: the
// three variance rounds with the single-pair kernels' arithmetic and trees
// (the product truncated to int per point, a 1024-wide FP32 tree per block of
// 1024 points, the block sums truncated to int and reduced by the same tree,
// clamped to at least n), so the sums are the same bit for bit. var: three
// values in LDS, unchanged when n = 0 (no solve follows then).
template <int BS>
__device__ void student_t_wg(const float* __restrict__ r2, int n, float* var) {
    __shared__ float p[1024];
    __shared__ float bsum[512];  // n <= W*H <= 512 * 1024 (checked at preparation)
    const int t = threadIdx.x;
    if (n <= 0) return;
    const int nb = (n + 1023) / 1024;
    auto tree = [&]() {
        for (int s2 = 512; s2 >= 1; s2 >>= 1) {
            for (int e = t; e < s2; e += BS) p[e] += p[e + s2];
            __syncthreads();
        }
    };
    for (int round = 1; round <= 3; round++) {
        const float prev = round > 1 ? var[round - 2] : 0.0f;
        for (int bk = 0; bk < nb; bk++) {
            for (int e = t; e < 1024; e += BS) {
                const int i = bk * 1024 + e;
                int sum = 0;
                if (i < n) {
                    const float x = r2[i];
                    const float w = round == 1 ? 6 / (5 + x / 100.0f) : 6 / (5 + x / (prev / (float)n));
                    sum += x * w;
                }
                p[e] = (float)sum;
            }
            __syncthreads();
            tree();
            if (t == 0) bsum[bk] = p[0];
            __syncthreads();
        }
        if (nb == 1) {
            if (t == 0) var[round - 1] = bsum[0];
        } else {
            for (int e = t; e < 1024; e += BS) {
                int ss = 0;
                for (int q = e; q < nb; q += 1024) ss += bsum[q];
                p[e] = (float)ss;
            }
            __syncthreads();
            tree();
            if (t == 0) {
                float v = p[0];
                const float mn = (float)n;
                if (v < mn) v = mn;
                var[round - 1] = v;
            }
        }
        __syncthreads();
    }
}
// The geometric system of one iteration (icpResidualIndexKernel2,
// rgb2gray.cu:323-390, and the 27 dot products): per thread in FP32 over its
// points, per wave by halving exchanges, across waves in FP64 into sysg[27].
template <int BS>
__device__ void icp_system_wg(const PairDesc& d, int layer, int n, float v3, const float* T, const float* c,
                              double (*wsum)[29], double* sysg) {
    constexpr int NW = BS / 64;
    const int t = threadIdx.x, wave = t >> 6, lane = lane_id();
    const int w = d.W >> layer, h = d.H >> layer;
    const float a = 1.0f / (float)(1 << layer), b = 0.5f * (a - 1.0f);
    const RT_GLOBAL int* isel = gmem(d.isel);
    const RT_GLOBAL float* ir2 = gmem(d.ir2);
    float acc[27];
#pragma unroll
    for (int k = 0; k < 27; k++) acc[k] = 0.0f;
    for (int i = t; i < n; i += BS) {
        float res, J[6];
        icp_point_terms(d.rpos[layer], d.rnrm[layer], d.cpos[layer], w, h, isel[i], ir2[i], v3, n, T, c, a, b, 1e-3f,
                        res, J);
        int q = 0;
#pragma unroll
        for (int r0 = 0; r0 < 6; r0++)
#pragma unroll
            for (int c0 = r0; c0 < 6; c0++) acc[q++] += J[r0] * J[c0];
#pragma unroll
        for (int r0 = 0; r0 < 6; r0++) acc[21 + r0] += J[r0] * res;
    }
    float v32[32];
#pragma unroll
    for (int k = 0; k < 32; k++) v32[k] = k < 27 ? acc[k] : 0.0f;
    const float tot = wave_sum32(v32);
    if ((lane & 1) == 0 && (lane >> 1) < 27) wsum[wave][lane >> 1] = (double)tot;
    __syncthreads();
    if (t < 27) {
        double v = 0.0;
        for (int w2 = 0; w2 < NW; w2++) v += wsum[w2][t];
        sysg[t] = v;
    }
    __syncthreads();
}

// ------------------------------------------------------------------ persistent tracking
struct TrackCfg {
    int solver;
    int robust;
    int mad_bins;
    float scale_out;
    float lambda;  // bi-objective weight of the geometric system (linearFuse)
    int pc_cap;  // point-cache capacity (points; PC variants)
    int rz_len;  // length of the RZ residual array (the batch's largest point count)
};


// RZ: residual and zWeight of each point kept in registers (0) or in LDS (1,
// at most RZ_MAX points per pair), trading LDS for occupancy.
constexpr int RZ_MAX = 8192;
// PC: the first cfg.pc_cap points of a pair keep {x, y, z, reference gray of
// the level} in LDS across iterations (after the RZ residuals, when both are
// on). The record is filled by each level's first residual pass (xyz by the
// pair's first pass), so later iterations read 16 B of LDS instead of 20 B of
// point stream from HBM; points beyond the capacity stream as before. Each
// thread reads and writes only the entries of its own points, so the cache
// needs no barrier of its own.
constexpr int PC_MAX = 8192;
// WPE: waves per SIMD the register allocation must allow (1: no constraint
// beyond the workgroup size). Two 512-thread workgroups per CU (WPE 4), which
// would overlap one pair's serial phases with the other's point passes, were
// measured slower than one 1024-thread workgroup at 640x480: the kernel is
// bound by memory traffic, not by those phases.
// MODE: the tracking mode (RT_GEOMETRIC 0, RT_PHOTOMETRIC 1, RT_BIOBJECTIVE 2).
// The geometric modes add, per level, the ICP selection and Student-t weights
// at the level's first iteration and the geometric system every iteration
// (ShapeTracker::updateLinearSystem, shapetracker.cpp:357-382), fused with the
// photometric one by linearFuse in the bi-objective mode (:392-394).
template <int BS, int PPT, int RZ, int WPE = 1, int PC = 0, int MODE = 1>
__global__ __launch_bounds__(BS, WPE) void kb_track(const PairDesc* __restrict__ P, const int* __restrict__ order,
                                                    int npairs, int* queue, TrackCfg cfg) {
    constexpr int NW = BS / 64;
    // system terms: photometric 27 | geometric 27 (geometric modes) | the two
    // sums of the photometric residual RMS
    constexpr int NSYS = MODE == 1 ? 29 : 56, RMS0 = MODE == 1 ? 27 : 54;
    __shared__ float cal[96];
    __shared__ float T[16];
    __shared__ unsigned hist[256];
    __shared__ double wsum[NW][29];
    __shared__ double sys[NSYS];
    __shared__ float s_var[3];
    __shared__ double ws[SOLVE_WS];
    __shared__ int s_pair, s_med;
    __shared__ int s_ok;
    __shared__ float lut[256];
    // RZ: dynamic LDS after every static array, so that the static arrays keep
    // small addresses (immediate offsets from one base register)
    extern __shared__ float4 dyn_lds[];
    float2* rz_lds = reinterpret_cast<float2*>(dyn_lds);
    float4* pcache = dyn_lds + (RZ ? cfg.rz_len / 2 : 0);
    const int t = threadIdx.x, wave = t >> 6, lane = lane_id();
    load_lut(lut);
    for (;;) {
        __syncthreads();
        if (t == 0) {
            int q = atomicAdd(queue, 1);
            s_pair = q < npairs ? order[q] : -1;
        }
        __syncthreads();
        const int p = __builtin_amdgcn_readfirstlane(s_pair);
        if (p < 0) break;
        const PairDesc& d = P[p];
        const int W = d.W, H = d.H, L = d.L, N = d.N;
        const int C = PC ? min(N, cfg.pc_cap) : 0;  // points held in the cache
        bool xyz_cached = false;                 // set after the pair's first residual pass
        if (t < 96) cal[t] = d.calib[t];
        if (t < 16) T[t] = d.trel_in[t];
        if (t == 0) s_ok = 1;
        __syncthreads();
        // calibration and pose are the same for every lane: scalar registers
        float cl[96];
#pragma unroll
        for (int k = 0; k < 96; k++) cl[k] = uni(cal[k]);
        const bool simple = d.simple != 0;
        float rms_last = 0.0f;
        for (int layer = L - 1; layer >= 0; layer--) {
            const int nit = d.iters[L - 1 - layer];
            const float* img = d.gray[1][layer];
            const JacView jv = jac_view(d.jac[layer], N);
            const int wl = W >> layer, hl = H >> layer;
            const float a = 1.0f / (float)(1 << layer), b = 0.5f * (a - 1.0f);
            int nicp = 0;  // ICP points of the level (geometric modes)
            for (int it = 0; it < nit; it++) {
                for (int k = t; k < 256; k += BS) hist[k] = 0u;
                float Tl[16];
#pragma unroll
                for (int k = 0; k < 16; k++) Tl[k] = uni(T[k]);
                __syncthreads();
              if constexpr (MODE != 0) {
                // residual pass: r and zWeight stay in registers (RZ 0) or go to LDS (RZ 1)
                float r[RZ ? 1 : PPT], zw[RZ ? 1 : PPT];
                auto residual = [&](int i, float& rr, float& zz) {
                    float4 p4;
                    float ir;
                    if (PC && i < C) {
                        if (it == 0) {
                            // the level's first pass: xyz from HBM on the pair's first pass,
                            // else from the cache; the level's reference gray from HBM
                            p4 = xyz_cached ? pcache[i] : ldg(d.pts, i);
                            ir = ldg(jv.iref, i);
                            pcache[i] = make_float4(p4.x, p4.y, p4.z, ir);
                        } else {
                            p4 = pcache[i];
                            ir = p4.w;
                        }
                    } else {
                        p4 = ldg(d.pts, i);
                        ir = ldg(jv.iref, i);
                    }
                    const U16BlockDepthSrc zs{gmem(d.zblk)};  // built by kb_pyramid
                    if (layer == 0) {
                        const U8Src g{d.g8[1], lut};
                        if (simple) photo_point<true>(p4, ir, Tl, a, b, g, wl, hl, zs, W, H, cl, rr, zz);
                        else photo_point<false>(p4, ir, Tl, a, b, g, wl, hl, zs, W, H, cl, rr, zz);
                    } else {
                        const F32Src g{img};
                        if (simple) photo_point<true>(p4, ir, Tl, a, b, g, wl, hl, zs, W, H, cl, rr, zz);
                        else photo_point<false>(p4, ir, Tl, a, b, g, wl, hl, zs, W, H, cl, rr, zz);
                    }
                    int bn = (int)((float)(cfg.mad_bins - 1) * fabsf(rr));
                    if (bn >= 0 && bn < cfg.mad_bins) atomicAdd(&hist[bn], 1u);
                };
                if constexpr (RZ) {
#pragma unroll PPT
                    for (int i = t; i < N; i += BS) {
                        float rr, zz;
                        residual(i, rr, zz);
                        rz_lds[i] = make_float2(rr, zz);
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < PPT; k++) {
                        int i = t + k * BS;
                        r[k] = 0.0f; zw[k] = 0.0f;
                        if (i < N) residual(i, r[k], zw[k]);
                    }
                }
                xyz_cached = true;
                __syncthreads();
                // median bin: first bin whose running count exceeds N/2 (histogram64.cu:48-61)
                if (wave == 0) {
                    unsigned c4[4];
                    unsigned run = 0;
#pragma unroll
                    for (int q = 0; q < 4; q++) { c4[q] = hist[lane * 4 + q]; run += c4[q]; }
                    unsigned incl = run;
#pragma unroll
                    for (int o = 1; o < 64; o <<= 1) {
                        unsigned v = __shfl_up(incl, o, 64);
                        if (lane >= o) incl += v;
                    }
                    unsigned acc = incl - run;
                    int found = 1 << 30;
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        acc += c4[q];
                        if (found == (1 << 30) && (int)acc > N / 2) found = lane * 4 + q;
                    }
                    int m = found;
#pragma unroll
                    for (int o = 32; o >= 1; o >>= 1) m = min(m, __shfl_xor(m, o, 64));
                    if (lane == 0) s_med = m == (1 << 30) ? 0 : m;
                }
                __syncthreads();
                const float s = mad_scale(s_med, cfg.mad_bins);
                float acc[29];
#pragma unroll
                for (int k = 0; k < 29; k++) acc[k] = 0.0f;
                auto accumulate = [&](int i, float rk, float zk) {
                    float w = robust_weight(rk, zk, s, cfg.robust);
                    float wr = rk * w;
                    float4 ja = ldg(jv.a, i);
                    float2 jb = ldg(jv.b, i);
                    float jw[6] = {ja.x * w, ja.y * w, ja.z * w, ja.w * w, jb.x * w, jb.y * w};
                    // each lane's running sums by fused multiply-adds (one rounding
                    // per term: these sums are this kernel's own, in their own
                    // order, and agree with the oracle's to rounding)
                    int q = 0;
#pragma unroll
                    for (int r0 = 0; r0 < 6; r0++)
#pragma unroll
                        for (int c0 = r0; c0 < 6; c0++) { acc[q] = __builtin_fmaf(jw[r0], jw[c0], acc[q]); q++; }
#pragma unroll
                    for (int r0 = 0; r0 < 6; r0++) acc[21 + r0] = __builtin_fmaf(jw[r0], wr, acc[21 + r0]);
                    acc[27] = __builtin_fmaf(wr, wr, acc[27]);
                    acc[28] = __builtin_fmaf(w, w, acc[28]);
                };
                if constexpr (RZ) {
#pragma unroll PPT
                    for (int i = t; i < N; i += BS) {
                        float2 v = rz_lds[i];
                        accumulate(i, v.x, v.y);
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < PPT; k++) {
                        int i = t + k * BS;
                        if (i < N) accumulate(i, r[k], zw[k]);
                    }
                }
                {
                    float v32[32];
#pragma unroll
                    for (int k = 0; k < 32; k++) v32[k] = k < 29 ? acc[k] : 0.0f;
                    const float tot = wave_sum32(v32);
                    if ((lane & 1) == 0 && (lane >> 1) < 29) wsum[wave][lane >> 1] = (double)tot;
                }
                __syncthreads();
                // wave sums combined in FP64 in wave order, one term per lane
                if (t < 29) {
                    double v = 0.0;
                    for (int w2 = 0; w2 < NW; w2++) v += wsum[w2][t];
                    sys[t < 27 ? t : RMS0 + t - 27] = v;
                }
                __syncthreads();
              }
                if constexpr (MODE != 1) {
                    // the level's ICP points and Student-t weights at its first
                    // iteration (under the pose the iteration starts from)
                    if (it == 0) {
                        nicp = icp_select_wg<BS>(d, layer, Tl, cl);
                        student_t_wg<BS>(d.ir2, nicp, s_var);
                        if (t == 0) {
                            d.icnt[layer] = nicp;
                            for (int k = 0; k < 3; k++) d.ivar[layer * 3 + k] = nicp > 0 ? s_var[k] : 0.0f;
                        }
                    }
                    icp_system_wg<BS>(d, layer, nicp, s_var[2], Tl, cl, wsum, sys + 27);
                }
                if (wave == 0) {
                    SolveCfg sc;
                    sc.solver = cfg.solver; sc.mode = MODE; sc.lambda = cfg.lambda; sc.scale_out = cfg.scale_out;
                    sc.robust = cfg.robust; sc.mad_bins = cfg.mad_bins; sc.scale_in = 1e-3f;
                    // no update from an empty geometric selection (as the single-pair path)
                    if (MODE == 1 || nicp > 0) {
                        const bool ok = solve_update_wave(sys, T, sc, N, nicp, ws);
                        if (t == 0 && !ok) s_ok = 0;
                    }
                    if (t == 0 && MODE != 0)
                        rms_last = sys[RMS0 + 1] > 0.0 ? (float)sqrt(sys[RMS0] / sys[RMS0 + 1]) : 0.0f;
                }
                __syncthreads();
            }
        }
        if (t < 16) d.trel_out[t] = T[t];
        if (t == 0) { *d.rms = rms_last; *d.status = s_ok ? 0 : -1; }
    }
}

typedef void (*TrackFn)(const PairDesc*, const int*, int, int*, TrackCfg);
struct Variant { int bs, ppt, rz, wpe; TrackFn fn; int pc = 0; int mode = RT_PHOTOMETRIC; };
static const Variant kVariants[] = {
    {256, 8, 0, 1, kb_track<256, 8, 0>},     {256, 16, 0, 1, kb_track<256, 16, 0>},
    {256, 32, 0, 1, kb_track<256, 32, 0>},   {512, 4, 0, 1, kb_track<512, 4, 0>},
    {512, 8, 0, 1, kb_track<512, 8, 0>},     {512, 16, 0, 1, kb_track<512, 16, 0>},
    {512, 32, 0, 1, kb_track<512, 32, 0>},   {1024, 4, 0, 1, kb_track<1024, 4, 0>},
    {1024, 8, 0, 1, kb_track<1024, 8, 0>},   {1024, 16, 0, 1, kb_track<1024, 16, 0>},
    {1024, 32, 0, 1, kb_track<1024, 32, 0>},
    // RZ variants: ppt is the unroll depth of a loop over the pair's points
    {256, 2, 1, 1, kb_track<256, 2, 1>},     {256, 4, 1, 1, kb_track<256, 4, 1>},
    {512, 2, 1, 1, kb_track<512, 2, 1>},     {512, 4, 1, 1, kb_track<512, 4, 1>},
    {1024, 2, 1, 1, kb_track<1024, 2, 1>},
    // point-cache variants; pc = capacity in points (16 B of LDS each). With LDS
    // residuals, 5632 points fill the 160 KB of LDS: 8 KB static + 64 KB of
    // residuals for 8192 points + 88 KB of cache.
    {1024, 8, 0, 1, kb_track<1024, 8, 0, 1, 1>, 8192},  {1024, 16, 0, 1, kb_track<1024, 16, 0, 1, 1>, 8192},
    {512, 16, 0, 1, kb_track<512, 16, 0, 1, 1>, 8192},
    {1024, 2, 1, 1, kb_track<1024, 2, 1, 1, 1>, 5632},   {1024, 4, 1, 1, kb_track<1024, 4, 1, 1, 1>, 5632},
    {512, 4, 1, 1, kb_track<512, 4, 1, 1, 1>, 5632},     {1024, 2, 1, 1, kb_track<1024, 2, 1, 1, 1>, 5120},
    // bi-objective: the photometric variants' shapes with the geometric terms
    // (the selection's LDS leaves room for a smaller point cache)
    {1024, 2, 1, 1, kb_track<1024, 2, 1, 1, 1, RT_BIOBJECTIVE>, 4608, RT_BIOBJECTIVE},
    {1024, 4, 1, 1, kb_track<1024, 4, 1, 1, 1, RT_BIOBJECTIVE>, 4608, RT_BIOBJECTIVE},
    {1024, 2, 1, 1, kb_track<1024, 2, 1, 1, 0, RT_BIOBJECTIVE>, 0, RT_BIOBJECTIVE},
    {1024, 8, 0, 1, kb_track<1024, 8, 0, 1, 0, RT_BIOBJECTIVE>, 0, RT_BIOBJECTIVE},
    {512, 4, 1, 1, kb_track<512, 4, 1, 1, 0, RT_BIOBJECTIVE>, 0, RT_BIOBJECTIVE},
    {256, 4, 1, 1, kb_track<256, 4, 1, 1, 0, RT_BIOBJECTIVE>, 0, RT_BIOBJECTIVE},
    // geometric only: no photometric points (ppt and rz unused)
    {1024, 1, 0, 1, kb_track<1024, 1, 0, 1, 0, RT_GEOMETRIC>, 0, RT_GEOMETRIC},
    {512, 1, 0, 1, kb_track<512, 1, 0, 1, 0, RT_GEOMETRIC>, 0, RT_GEOMETRIC},
    {256, 1, 0, 1, kb_track<256, 1, 0, 1, 0, RT_GEOMETRIC>, 0, RT_GEOMETRIC},
};

}  // namespace rt

using namespace rt;

// ------------------------------------------------------------------ C ABI
struct rt_batch {
    int device = 0, max_pairs = 0, max_w = 0, max_h = 0;
    rt_params prm;
    hipStream_t stream = nullptr;
    // A second stream beside the keyframe preparation: the current frame's
    // pyramid (pyramid: its end) and the pose network's proposals (they read
    // only the frames and write the initial poses; proposed: their end). It
    // waits for the prepare call's uploads (inputs), and the proposals for the
    // end of the gradient bins (binned); the batch stream waits for both
    // before tracking.
    hipStream_t aux = nullptr;
    hipEvent_t inputs = nullptr, proposed = nullptr, pyramid = nullptr, binned = nullptr;
    // The pose network's inputs, built by the preparation from the frames it
    // reads anyway once rt_batch_propose has been called on this batch
    // (feat_on): max_pairs x (max_w / 8) x (max_h / 8) x 4 bfloat16;
    // feat_ready: the last preparation built them for every pair.
    bf16* feat = nullptr;
    bool feat_on = false, feat_ready = false;
    int cus = 256;
    std::vector<void*> allocs;
    // per-pair arenas
    char* arena = nullptr;
    size_t pair_bytes = 0;
    PairDesc* desc_dev = nullptr;
    int* order_dev = nullptr;
    int* queue = nullptr;
    float* calib_dev = nullptr;   // 96 per pair
    float* trel_in = nullptr;     // 16 per pair
    float* trel_out = nullptr;
    float* rms = nullptr;
    int* status = nullptr;
    // pinned staging of the host tables each prepare / retrack uploads (the
    // copies are queued, so their sources must outlive the call)
    HostStage st_desc, st_calib, st_order, st_trel;
    HostStage st_res;  // page-locked landing buffer of the results (poses, rms, status)
    // last batch
    int n = 0;
    std::vector<PairDesc> desc, desc_next;  // the prepared pairs; the next preparation's, swapped in on success
    int maxW = 0, maxH = 0, maxN = 0;
    // tuning and graphs
    std::map<std::string, std::pair<int, int>> tuned;  // key -> (variant, grid)
    int variant = -1, grid = 0;
    hipGraphExec_t graph_pyr = nullptr, graph_bins = nullptr, graph_sel = nullptr, graph_track = nullptr;
    int graph_n = -1, graph_variant = -1, graph_grid = -1, graph_maxN = -1;  // the track graph's launch shape
};

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Device memory a batch's descriptors point into: per-pair arenas of
// pair_bytes (sized for max_w x max_h and nsel points) and the per-pair tables.
struct BatchMem {
    char* arena;
    size_t pair_bytes;
    int max_w, max_h;
    size_t nsel;
    bool icp;         // geometric buffers in the arena (geometric and bi-objective modes)
    float* calib;     // 96 per pair
    float* trel_in;   // 16 per pair
    float* trel_out;  // 16 per pair
    float* rms;
    int* status;
};
static size_t pair_arena_bytes(int max_w, int max_h, size_t N, bool icp = false) {
    const size_t npx = (size_t)max_w * max_h;
    size_t per = 0;  // inputs are borrowed device pointers; the arena holds derived data only
    per += 2 * align256(npx * 4 / 3 + 64);      // two gray pyramids, levels 1..3
    per += align256(npx);                       // gradient-magnitude bins
    per += align256((size_t)max_w * ((max_h + 7) & ~7) * 2);  // current depth map in 8 x 8 blocks
    per += align256(256 * 4) + align256(N * 4) + align256(N * 16) + 4 * align256(jac_bytes((int)N));
    if (icp) {
        for (int l = 0; l < 4; l++) per += 3 * align256((npx >> (2 * l)) * 16);  // rpos, rnrm, cpos
        per += 2 * align256(npx * 4) + align256(64);                                // isel, ir2, statistics
    }
    return per;
}


extern "C" int rt_batch_create(int device, int max_pairs, int max_w, int max_h, const rt_params* params,
                               rt_batch** out) {
    if (!out || max_pairs < 1 || max_w < 32 || max_h < 32) return set_error(RT_ERR_ARG, "rt_batch_create: bad argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return set_error(RT_ERR_ARG, "no such device");
    if (hipSetDevice(device) != hipSuccess) return set_error(RT_ERR_HIP, "hipSetDevice");
    rt_batch* b = new rt_batch();
    b->device = device; b->max_pairs = max_pairs; b->max_w = max_w; b->max_h = max_h;
    if (params) b->prm = *params; else rt_default_params(&b->prm);
    if (b->prm.n_select < 256 || b->prm.n_select % 256 || b->prm.n_select > 32768) { delete b; return set_error(RT_ERR_ARG, "n_select must be a multiple of 256 up to 32768 for batches"); }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) b->cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) { delete b; return set_error(RT_ERR_HIP, "stream"); }
    if (hipStreamCreateWithFlags(&b->aux, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->inputs, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->proposed, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->pyramid, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->binned, hipEventDisableTiming) != hipSuccess) {
        rt_batch_destroy(b);
        return set_error(RT_ERR_HIP, "stream");
    }
    if (b->prm.mode < RT_GEOMETRIC || b->prm.mode > RT_BIOBJECTIVE) { rt_batch_destroy(b); return set_error(RT_ERR_ARG, "bad mode"); }
    b->prm.lambda = std::min(1.0f, std::max(0.0f, b->prm.lambda));  // setLambda's clamp (phototracker.cpp:163-168)
    b->pair_bytes = pair_arena_bytes(max_w, max_h, (size_t)b->prm.n_select, b->prm.mode != RT_PHOTOMETRIC);
    size_t total = b->pair_bytes * (size_t)max_pairs;
    void* q = nullptr;
    if (hipMalloc(&q, total) != hipSuccess) { rt_batch_destroy(b); return set_error(RT_ERR_NOMEM, "batch arena"); }
    b->arena = (char*)q; b->allocs.push_back(q);
    auto dal = [&](void** p, size_t bytes) {
        if (hipMalloc(p, bytes) != hipSuccess) return false;
        hipMemset(*p, 0, bytes);
        b->allocs.push_back(*p);
        return true;
    };
    bool ok = dal((void**)&b->desc_dev, sizeof(PairDesc) * max_pairs) && dal((void**)&b->order_dev, 4 * max_pairs) &&
              dal((void**)&b->queue, 256) && dal((void**)&b->calib_dev, 96 * 4 * (size_t)max_pairs) &&
              dal((void**)&b->trel_in, 64 * (size_t)max_pairs) && dal((void**)&b->trel_out, 64 * (size_t)max_pairs) &&
              dal((void**)&b->rms, 4 * (size_t)max_pairs) && dal((void**)&b->status, 4 * (size_t)max_pairs);
    if (!ok) {
        rt_batch_destroy(b);
        return set_error(RT_ERR_NOMEM, "batch buffers");
    }
    *out = b;
    return RT_OK;
}

extern "C" void rt_batch_destroy(rt_batch* b) {
    if (!b) return;
    hipSetDevice(b->device);
    if (b->aux) hipStreamSynchronize(b->aux);
    if (b->stream) hipStreamSynchronize(b->stream);
    if (b->graph_pyr) hipGraphExecDestroy(b->graph_pyr);
    if (b->graph_bins) hipGraphExecDestroy(b->graph_bins);
    if (b->graph_sel) hipGraphExecDestroy(b->graph_sel);
    if (b->graph_track) hipGraphExecDestroy(b->graph_track);
    for (void* p : b->allocs) hipFree(p);
    b->st_desc.release(); b->st_calib.release(); b->st_order.release(); b->st_trel.release(); b->st_res.release();
    if (b->inputs) hipEventDestroy(b->inputs);
    if (b->proposed) hipEventDestroy(b->proposed);
    if (b->pyramid) hipEventDestroy(b->pyramid);
    if (b->binned) hipEventDestroy(b->binned);
    if (b->feat) hipFree(b->feat);
    if (b->aux) hipStreamDestroy(b->aux);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
}

extern "C" void* rt_batch_stream(rt_batch* b) { return b ? (void*)b->stream : nullptr; }

// Where the launches of one batch find their tables: the batch's own buffers,
// or the fixed scratch of rt_batch_track_noalloc.
struct BatchView {
    const PairDesc* desc;
    const int* order;
    int* queue;
    int n, maxW, maxH, maxN;
    int solver, robust, mad_bins;
    int mode;
    float lambda;
};
static BatchView view_of(const rt_batch* b) {
    return BatchView{b->desc_dev, b->order_dev, b->queue, b->n, b->maxW, b->maxH, b->maxN,
                     b->prm.solver, b->prm.robust, b->prm.mad_bins, b->prm.mode, b->prm.lambda};
}

// Keyframe preparation in three parts: the current frame's pyramid (part 1),
// the histogram clear and the gradient bins with the keyframe's pyramid (part
// 2), the selection, points and clouds (part 3); 0: all.
static void launch_prep(const BatchView& v, hipStream_t s, int part = 0) {
    const int n = v.n;
    if (part == 0 || part == 1) {
        dim3 g64((v.maxW + 63) / 64, (v.maxH + 63) / 64, n);
        kb_pyramid<<<g64, 256, 0, s>>>(v.desc, 1);
    }
    if (part == 1) return;
    if (part != 3) {
        kb_clear_hist<<<n, 256, 0, s>>>(v.desc);
        dim3 gb((v.maxW + 63) / 64, (v.maxH + BINS_TILE_H - 1) / BINS_TILE_H, n);
        kb_bins<<<gb, 256, 0, s>>>(v.desc);
    }
    if (part == 2) return;
    kb_select<<<n, SEL_BS, 0, s>>>(v.desc);
    const int chunks = (v.maxN + 255) / 256;
    kb_ref_points<<<8 * ((n + 7) / 8) * chunks, 256, 0, s>>>(v.desc, n, chunks);
    if (v.mode != RT_PHOTOMETRIC) {
        kb_clouds<<<dim3((v.maxW + 15) / 16, (v.maxH + 15) / 16, n), 256, 0, s>>>(v.desc);
        for (int l = 1; l < 4; l++)
            kb_clouds_down<<<dim3(((v.maxW >> l) * (v.maxH >> l) + 255) / 256, n), 256, 0, s>>>(v.desc, l);
    }
}

static TrackCfg track_cfg(const BatchView& bv, const Variant& v) {
    TrackCfg c;
    c.lambda = bv.lambda;
    c.pc_cap = v.pc;
    c.rz_len = bv.maxN;
    c.solver = bv.solver;
    c.robust = bv.robust;
    c.mad_bins = bv.mad_bins;
    c.scale_out = 1e3f;
    return c;
}

static size_t track_dyn_lds(const Variant& v, int maxN) {
    size_t n = 0;
    if (v.rz && v.mode != RT_GEOMETRIC) n += sizeof(float2) * (size_t)maxN;
    if (v.pc) n += sizeof(float4) * (size_t)std::min(maxN, v.pc);
    return n;
}

constexpr size_t LDS_PER_CU = 160 * 1024;  // gfx950
constexpr int NUM_VARIANTS = (int)(sizeof(kVariants) / sizeof(kVariants[0]));
static_assert(NUM_VARIANTS <= 64, "variant tables below hold 64 entries");

// RZ and point-cache variants need more dynamic LDS than the default limit
// allows. A variant is usable only when its static LDS plus the largest dynamic
// LDS it may be launched with fits the CU's 160 KiB: a launch beyond that is
// rejected and would otherwise time as ~0 ms in the autotuner.
static bool lds_variant_ready(const Variant& v, int maxN) {
    static bool tried[64] = {}, ok[64] = {};
    int idx = (int)(&v - kVariants);
    if (!tried[idx]) {
        tried[idx] = true;
        const size_t need = (v.rz ? sizeof(float2) * RZ_MAX : 0) + sizeof(float4) * v.pc;
        hipFuncAttributes fa;
        ok[idx] = hipFuncGetAttributes(&fa, (const void*)v.fn) == hipSuccess && fa.sharedSizeBytes + need <= LDS_PER_CU &&
                  (need == 0 || hipFuncSetAttribute((const void*)v.fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                    (int)need) == hipSuccess);
        (void)hipGetLastError();
    }
    return ok[idx] && (!v.rz || maxN <= RZ_MAX);
}

static void launch_track(const BatchView& bv, int variant, int grid, hipStream_t s) {
    kb_reset_queue<<<1, 64, 0, s>>>(bv.queue);
    const Variant& v = kVariants[variant];
    v.fn<<<grid, v.bs, track_dyn_lds(v, bv.maxN), s>>>(bv.desc, bv.order, bv.n, bv.queue, track_cfg(bv, v));
}
static void launch_track(rt_batch* b, int variant, int grid, hipStream_t s) { launch_track(view_of(b), variant, grid, s); }

static std::string tune_key(const rt_batch* b) {
    // batch size bucket and the size mix decide the best configuration
    std::map<int, int> mix;
    for (const PairDesc& d : b->desc) mix[d.W * 10000 + d.H]++;
    char buf[64];
    std::string k = std::to_string(b->n) + ":" + std::to_string(b->maxN);
    for (auto& kv : mix) { snprintf(buf, sizeof(buf), "|%d*%d", kv.first, kv.second); k += buf; }
    return k;
}

// Time every variant that covers the largest point count, at 1, 2 and 4
// workgroups per CU, on the prepared batch; keep the fastest (per batch key).
static int autotune(rt_batch* b) {
    std::string key = tune_key(b);
    auto it = b->tuned.find(key);
    if (it != b->tuned.end()) { b->variant = it->second.first; b->grid = it->second.second; return RT_OK; }
    const char* force = getenv("RT_BATCH_VARIANT");
    const int verbose = getenv("RT_VERBOSE") ? atoi(getenv("RT_VERBOSE")) : 0;
    hipEvent_t e0, e1;
    hipEventCreate(&e0); hipEventCreate(&e1);
    float best = 1e30f;
    int bv = -1, bg = 0;
    for (int v = 0; v < NUM_VARIANTS; v++) {
        const Variant& var = kVariants[v];
        if (var.mode != b->prm.mode) continue;
        if (var.mode == RT_GEOMETRIC) {
            if (!lds_variant_ready(var, b->maxN)) continue;  // no photometric points to cover
        } else if (var.rz) {
            if (!lds_variant_ready(var, b->maxN)) continue;
        } else {
            if (var.bs * var.ppt < b->maxN) continue;
            if (var.bs * var.ppt >= 4 * b->maxN && var.ppt > 4) continue;  // mostly idle lanes
            if (var.pc && !lds_variant_ready(var, b->maxN)) continue;
        }
        if (force && atoi(force) != v) continue;
        // waves per CU stay within the hardware limit; a point-cache workgroup fills the CU's LDS
        int per_cu_max = 2048 / var.bs;
        for (int k = 1; k <= 4 && k <= per_cu_max; k *= 2) {
            int grid = std::min(b->n, b->cus * k);
            (void)hipGetLastError();
            launch_track(b, v, grid, b->stream);  // warm
            if (hipGetLastError() != hipSuccess) break;  // launch rejected: not a candidate
            hipEventRecord(e0, b->stream);
            launch_track(b, v, grid, b->stream);
            hipEventRecord(e1, b->stream);
            if (hipEventSynchronize(e1) != hipSuccess) { hipEventDestroy(e0); hipEventDestroy(e1); return set_error(RT_ERR_HIP, "autotune run failed"); }
            float ms = 0;
            hipEventElapsedTime(&ms, e0, e1);
            if (verbose >= 2)
                fprintf(stderr, "rt_batch autotune   bs %d ppt %d rz %d pc %d grid %d: %.3f ms\n", var.bs, var.ppt,
                        var.rz, var.pc, grid, ms);
            if (ms < best) { best = ms; bv = v; bg = grid; }
            if (grid == b->n) break;
        }
    }
    hipEventDestroy(e0); hipEventDestroy(e1);
    if (bv < 0) return set_error(RT_ERR_UNSUPPORTED, "no batch kernel variant covers the point count");
    b->tuned[key] = {bv, bg};
    b->variant = bv; b->grid = bg;
    if (getenv("RT_VERBOSE"))
        fprintf(stderr, "rt_batch autotune %s -> bs %d ppt %d rz %d wpe %d pc %d grid %d (%.3f ms)\n", key.c_str(),
                kVariants[bv].bs, kVariants[bv].ppt, kVariants[bv].rz, kVariants[bv].wpe, kVariants[bv].pc, bg, best);
    return RT_OK;
}

// The kernel variant table, for tests that force each variant through
// RT_BATCH_VARIANT: info = {workgroup size, points per thread (unroll depth
// for rz), residuals in LDS, point-cache capacity, waves per SIMD, mode}.
extern "C" int rt_batch_variant_info(int v, int info[6]) {
    if (!info || v < 0 || v >= NUM_VARIANTS) return set_error(RT_ERR_ARG, "rt_batch_variant_info: no such variant");
    const Variant& k = kVariants[v];
    info[0] = k.bs; info[1] = k.ppt; info[2] = k.rz; info[3] = k.pc; info[4] = k.wpe; info[5] = k.mode;
    return RT_OK;
}

// prep: 0 the tracking launch, else the part of the preparation (launch_prep)
static int capture(rt_batch* b, hipGraphExec_t* out, int prep) {
    hipGraph_t g;
    if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return set_error(RT_ERR_HIP, "begin capture");
    if (prep) launch_prep(view_of(b), b->stream, prep);
    else launch_track(b, b->variant, b->grid, b->stream);
    if (hipStreamEndCapture(b->stream, &g) != hipSuccess) return set_error(RT_ERR_HIP, "end capture");
    if (*out) hipGraphExecDestroy(*out);
    *out = nullptr;
    hipError_t e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return set_error(RT_ERR_HIP, "graph instantiate");
    return RT_OK;
}

static int fetch_results(rt_batch* b, rt_pair_result* out) {
    // into page-locked memory: direct copies, no staging inside the runtime
    const size_t n = b->n;
    float* tr = (float*)b->st_res.get(24 * n * 4);
    if (!tr) return set_error(RT_ERR_HIP, "result buffer");
    float* rms = tr + 16 * n;
    int* st = (int*)(rms + n);
    if (hipMemcpyAsync(tr, b->trel_out, 16 * n * 4, hipMemcpyDeviceToHost, b->stream) != hipSuccess ||
        hipMemcpyAsync(rms, b->rms, n * 4, hipMemcpyDeviceToHost, b->stream) != hipSuccess ||
        hipMemcpyAsync(st, b->status, n * 4, hipMemcpyDeviceToHost, b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess)
        return set_error(RT_ERR_HIP, "result copy");
    for (int i = 0; i < b->n; i++) {
        double Td[16], Ti[16];
        for (int k = 0; k < 16; k++) Td[k] = tr[16 * (size_t)i + k];
        invert_rt4(Td, Ti);
        memcpy(out[i].Trel, Td, sizeof(Td));
        memcpy(out[i].T, Ti, sizeof(Ti));
        out[i].rms = rms[i];
        out[i].status = st[i];
    }
    return RT_OK;
}

static int run_track(rt_batch* b, rt_pair_result* out) {
    int rc = autotune(b);
    if (rc) return rc;
    // the captured launch depends on the batch size, the variant, the grid and,
    // through its dynamic LDS, on the largest point count
    if (!b->graph_track || b->graph_n != b->n || b->graph_variant != b->variant || b->graph_grid != b->grid ||
        b->graph_maxN != b->maxN) {
        if ((rc = capture(b, &b->graph_track, 0))) return rc;
        b->graph_n = b->n; b->graph_variant = b->variant; b->graph_grid = b->grid; b->graph_maxN = b->maxN;
    }
    if (hipGraphLaunch(b->graph_track, b->stream) != hipSuccess) return set_error(RT_ERR_HIP, "graph launch");
    return out ? fetch_results(b, out) : RT_OK;
}

// Checks the n pairs against the parameters and the memory's limits and fills
// their descriptors (desc), calibration blocks (calib, 96 per pair), initial
// poses (trel, 16 per pair) and the work order (decreasing cost); returns the
// largest width, height and point count.
static int build_batch(const rt_params& prm, const BatchMem& m, int n, const rt_pair* pairs, PairDesc* desc,
                       float* calib, float* trel, int* order, int& maxW, int& maxH, int& maxN) {
    maxW = maxH = maxN = 0;
    for (int i = 0; i < n; i++) {
        const rt_pair& p = pairs[i];
        const rt_params& q = prm;
        int L = q.levels > 0 ? q.levels : (p.width == 640 ? 4 : 3);
        int mm = std::max(4, 1 << (L - 1));
        if (!p.ref_gray || !p.ref_depth || !p.cur_gray || !p.cur_depth || !p.calib) return set_error(RT_ERR_ARG, "null pair input");
        if (p.width > m.max_w || p.height > m.max_h || p.width % (2 * mm) || p.height % mm || p.width < 32 || p.height < 32)
            return set_error(RT_ERR_ARG, "pair size not supported by this batch");
        if ((p.width * p.height) / 512 > 1024) return set_error(RT_ERR_ARG, "pair too large");
        int N = q.n_select;
        if (N > p.width * p.height / 2) return set_error(RT_ERR_ARG, "n_select above half the pixels of a pair");
        PairDesc& d = desc[i];
        memset(&d, 0, sizeof(d));
        d.W = p.width; d.H = p.height; d.L = L; d.N = N;
        for (int k = 0; k < 4; k++) d.iters[k] = q.iters[k];
        d.calib = m.calib + 96 * (size_t)i;
        memcpy(&calib[96 * (size_t)i], p.calib, 96 * 4);
        d.simple = calib_simple(p.calib) ? 1 : 0;
        d.ifx = 1.0f / p.calib[KL]; d.ify = 1.0f / p.calib[KL + 4];
        d.g8[0] = p.ref_gray; d.d16[0] = p.ref_depth; d.g8[1] = p.cur_gray; d.d16[1] = p.cur_depth;
        char* base = m.arena + m.pair_bytes * (size_t)i;
        size_t npx = (size_t)p.width * p.height, off = 0;
        for (int f = 0; f < 2; f++) {
            size_t lo = 0;
            d.gray[f][0] = nullptr;  // level 0 is read from the 8-bit input
            for (int l = 1; l < 4; l++) {
                d.gray[f][l] = l < L ? (float*)(base + off + lo) : nullptr;
                lo += (npx >> (2 * l)) * 4;
            }
            off += align256((size_t)m.max_w * m.max_h * 4 / 3 + 64);
        }
        d.z[0] = d.z[1] = nullptr;   // depth is read from the 16-bit input
        d.bins = (uint8_t*)(base + off); off += align256((size_t)m.max_w * m.max_h);
        d.zblk = (uint16_t*)(base + off); off += align256((size_t)m.max_w * ((m.max_h + 7) & ~7) * 2);
        d.ghist = (unsigned*)(base + off); off += align256(256 * 4);
        const size_t NN = m.nsel;
        d.sel = (int*)(base + off); off += align256(NN * 4);
        d.pts = (float4*)(base + off); off += align256(NN * 16);
        for (int l = 0; l < 4; l++) { d.jac[l] = (float4*)(base + off); off += align256(jac_bytes((int)NN)); }
        if (m.icp) {
            const size_t mpx = (size_t)m.max_w * m.max_h;
            for (int l = 0; l < 4; l++) {
                const size_t sz = align256((mpx >> (2 * l)) * 16);
                d.rpos[l] = (float4*)(base + off); off += sz;
                d.rnrm[l] = (float4*)(base + off); off += sz;
                d.cpos[l] = (float4*)(base + off); off += sz;
            }
            d.isel = (int*)(base + off); off += align256(mpx * 4);
            d.ir2 = (float*)(base + off); off += align256(mpx * 4);
            d.icnt = (int*)(base + off);
            d.ivar = (float*)(base + off + 16); off += align256(64);
        }
        d.trel_in = m.trel_in + 16 * (size_t)i;
        d.trel_out = m.trel_out + 16 * (size_t)i;
        d.rms = m.rms + i;
        d.status = m.status + i;
        for (int k = 0; k < 16; k++) trel[16 * (size_t)i + k] = p.init_T ? (float)p.init_T[k] : ((k % 5 == 0) ? 1.0f : 0.0f);
        maxW = std::max(maxW, d.W); maxH = std::max(maxH, d.H); maxN = std::max(maxN, d.N);
    }
    // work queue order: decreasing cost (points x iterations), ties by index
    // (each pair's cost once; a batch of equal costs keeps index order, as
    // the stable sort would)
    std::vector<long> cost(n);
    std::vector<int> ord(n);
    bool uniform = true;
    for (int i = 0; i < n; i++) {
        long c = 0;
        for (int k = 0; k < desc[i].L; k++) c += desc[i].iters[k];
        cost[i] = c * desc[i].N + (long)desc[i].W * desc[i].H / 8;
        uniform = uniform && cost[i] == cost[0];
        ord[i] = i;
    }
    if (!uniform) std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return cost[x] > cost[y]; });
    memcpy(order, ord.data(), 4 * (size_t)n);  // order may be page-locked staging memory: written once
    return RT_OK;
}

extern "C" int rt_batch_prepare(rt_batch* b, int n, const rt_pair* pairs) {
    if (!b || !pairs || n < 1 || n > b->max_pairs) return set_error(RT_ERR_ARG, "rt_batch_prepare: bad argument");
    if (hipSetDevice(b->device) != hipSuccess) return set_error(RT_ERR_HIP, "hipSetDevice");
    // descriptors into a reused host vector; calibration blocks, initial poses
    // and the work order straight into the pinned staging buffers they are
    // uploaded from (no zero-filled temporaries, no second copies)
    std::vector<PairDesc>& desc = b->desc_next;
    desc.resize(n);
    float* calib = (float*)b->st_calib.get(96 * 4 * (size_t)n);
    float* trel = (float*)b->st_trel.get(64 * (size_t)n);
    int* order = (int*)b->st_order.get(4 * (size_t)n);
    if (!calib || !trel || !order) return set_error(RT_ERR_HIP, "rt_batch_prepare: staging memory");
    const BatchMem m{b->arena, b->pair_bytes, b->max_w, b->max_h, (size_t)b->prm.n_select, b->prm.mode != RT_PHOTOMETRIC,
                     b->calib_dev, b->trel_in, b->trel_out, b->rms, b->status};
    int maxW = 0, maxH = 0, maxN = 0;
    int rc = build_batch(b->prm, m, n, pairs, desc.data(), calib, trel, order, maxW, maxH, maxN);
    if (rc) return rc;
    // the network's inputs in the preparation: pairs of one size, a multiple
    // of 8, with the current frame 2-byte (gray) and 4-byte (depth) aligned
    bool feat = b->feat_on;
    for (int i = 0; i < n && feat; i++)
        feat = desc[i].W == desc[0].W && desc[i].H == desc[0].H && desc[i].W % 8 == 0 && desc[i].H % 8 == 0 &&
               !((uintptr_t)desc[i].g8[1] & 1) && !((uintptr_t)desc[i].d16[1] & 3);
    for (int i = 0; i < n; i++)
        desc[i].feat = feat ? b->feat + (size_t)i * (desc[i].W / 8) * (desc[i].H / 8) * 4 : nullptr;
    b->feat_ready = feat;
    if (!b->st_desc.upload(b->desc_dev, desc.data(), sizeof(PairDesc) * n, b->stream) ||
        !b->st_calib.send(b->calib_dev, 96 * 4 * (size_t)n, b->stream) ||
        !b->st_order.send(b->order_dev, 4 * (size_t)n, b->stream) ||
        !b->st_trel.send(b->trel_in, 64 * (size_t)n, b->stream))
        return set_error(RT_ERR_HIP, "descriptor upload");
    bool changed = b->n != n || b->maxW != maxW || b->maxH != maxH || b->maxN != maxN;
    b->n = n; b->desc.swap(desc); b->maxW = maxW; b->maxH = maxH; b->maxN = maxN;
    if (changed || !b->graph_pyr || !b->graph_bins || !b->graph_sel) {
        if ((rc = capture(b, &b->graph_pyr, 1)) || (rc = capture(b, &b->graph_bins, 2)) || (rc = capture(b, &b->graph_sel, 3)))
            return rc;
    }
    // The current frame's pyramid runs on the second stream beside the
    // keyframe's gradient bins (it streams the frame, they are bound by
    // arithmetic). The pose network's proposals (rt_batch_propose) follow it
    // there once the bins are done (binned), beside the selection and the
    // reference points, which are bound by latency. The batch stream waits for
    // the pyramid before anything after the preparation.
    if (hipEventRecord(b->inputs, b->stream) != hipSuccess || hipStreamWaitEvent(b->aux, b->inputs, 0) != hipSuccess ||
        hipGraphLaunch(b->graph_pyr, b->aux) != hipSuccess || hipEventRecord(b->pyramid, b->aux) != hipSuccess ||
        hipGraphLaunch(b->graph_bins, b->stream) != hipSuccess || hipEventRecord(b->binned, b->stream) != hipSuccess ||
        hipGraphLaunch(b->graph_sel, b->stream) != hipSuccess || hipStreamWaitEvent(b->stream, b->pyramid, 0) != hipSuccess)
        return set_error(RT_ERR_HIP, "graph launch");
    return RT_OK;
}

extern "C" int rt_batch_track(rt_batch* b, int n, const rt_pair* pairs, rt_pair_result* out) {
    int rc = rt_batch_prepare(b, n, pairs);
    return rc ? rc : run_track(b, out);
}

extern "C" int rt_batch_propose(rt_batch* b, rt_cnn* net) {
    if (!b || !net || b->n < 1) return set_error(RT_ERR_STATE, "rt_batch_propose: no prepared batch or no network");
    if (hipSetDevice(b->device) != hipSuccess) return set_error(RT_ERR_HIP, "hipSetDevice");
    std::vector<FrameSrc> src(b->n);
    for (int i = 0; i < b->n; i++) {
        const PairDesc& d = b->desc[i];
        src[i].W = d.W; src[i].H = d.H;
        src[i].g8[0] = d.g8[0]; src[i].g8[1] = d.g8[1];
        src[i].d16[0] = d.d16[0]; src[i].d16[1] = d.d16[1];
    }
    // inputs built by the preparation when it built them for this network's
    // input size; else the network builds them (the same values)
    const bool ready = b->feat_ready && cnn_cells8(net, b->desc[0].W, b->desc[0].H);
    if (!b->feat_on) {
        // from the next preparation on: one buffer for the batch's lifetime
        const size_t bytes = (size_t)b->max_pairs * (b->max_w / 8) * (b->max_h / 8) * 4 * sizeof(bf16);
        if (hipMalloc(&b->feat, bytes) != hipSuccess) return set_error(RT_ERR_HIP, "rt_batch_propose: input buffer");
        b->feat_on = true;
    }
    if (hipStreamWaitEvent(b->aux, b->binned, 0) != hipSuccess) return set_error(RT_ERR_HIP, "rt_batch_propose: wait");
    // after the pyramid and the bins
    const int rc = cnn_propose(net, b->n, src.data(), b->trel_in, b->aux, ready ? b->feat : nullptr);
    if (hipEventRecord(b->proposed, b->aux) != hipSuccess || hipStreamWaitEvent(b->stream, b->proposed, 0) != hipSuccess)
        return set_error(RT_ERR_HIP, "rt_batch_propose: join");
    return rc;
}

// Inspection of one pair of the last prepared batch: RT_DBG_GRAD_HIST,
// RT_DBG_SELECTED, RT_DBG_POINTS, RT_DBG_JACOBIAN (level), RT_DBG_TREL (initial
// pose of the pair); in the geometric modes RT_DBG_REF_CLOUD, RT_DBG_CUR_CLOUD
// (level) and, after tracking, RT_DBG_ICP_COUNT (the four levels' counts) and
// RT_DBG_ICP_VAR (level: the three Student-t sums).
extern "C" int rt_batch_debug_read(rt_batch* b, int pair, int what, int level, void* host, size_t bytes) {
    if (!b || !host || pair < 0 || pair >= b->n) return set_error(RT_ERR_ARG, "rt_batch_debug_read: bad argument");
    if (hipSetDevice(b->device) != hipSuccess) return set_error(RT_ERR_HIP, "hipSetDevice");
    const PairDesc& d = b->desc[pair];
    const void* src = nullptr;
    size_t n = 0;
    switch (what) {
        case RT_DBG_GRAD_HIST: src = d.ghist; n = 256 * 4; break;
        case RT_DBG_SELECTED: src = d.sel; n = (size_t)d.N * 4; break;
        case RT_DBG_POINTS: src = d.pts; n = (size_t)d.N * 16; break;
        case RT_DBG_JACOBIAN:
            if (level < 0 || level >= d.L) return set_error(RT_ERR_ARG, "bad level");
            src = d.jac[level]; n = jac_bytes(d.N); break;
        case RT_DBG_TREL: src = d.trel_in; n = 64; break;
        case RT_DBG_DEPTH_BLOCKS: src = d.zblk; n = (size_t)d.W * ((d.H + 7) & ~7) * 2; break;
        case RT_DBG_REF_CLOUD:
        case RT_DBG_CUR_CLOUD:
        case RT_DBG_ICP_COUNT:
        case RT_DBG_ICP_VAR: {
            if (!d.isel) return set_error(RT_ERR_STATE, "rt_batch_debug_read: geometric buffers exist in the geometric modes only");
            if (level < 0 || level >= d.L) return set_error(RT_ERR_ARG, "bad level");
            const size_t nl = (size_t)(d.W >> level) * (d.H >> level);
            if (what == RT_DBG_REF_CLOUD) {
                if (bytes < nl * 32) return set_error(RT_ERR_ARG, "buffer too small");
                if (hipMemcpyAsync(host, d.rpos[level], nl * 16, hipMemcpyDeviceToHost, b->stream) != hipSuccess ||
                    hipMemcpyAsync((char*)host + nl * 16, d.rnrm[level], nl * 16, hipMemcpyDeviceToHost, b->stream) != hipSuccess ||
                    hipStreamSynchronize(b->stream) != hipSuccess)
                    return set_error(RT_ERR_HIP, "rt_batch_debug_read: copy");
                return RT_OK;
            }
            if (what == RT_DBG_CUR_CLOUD) { src = d.cpos[level]; n = nl * 16; }
            else if (what == RT_DBG_ICP_COUNT) { src = d.icnt; n = 16; }
            else { src = d.ivar + 3 * level; n = 12; }
            break;
        }
        default: return set_error(RT_ERR_UNSUPPORTED, "rt_batch_debug_read: not available for batches");
    }
    if (bytes < n) return set_error(RT_ERR_ARG, "buffer too small");
    if (hipMemcpyAsync(host, src, n, hipMemcpyDeviceToHost, b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess)
        return set_error(RT_ERR_HIP, "rt_batch_debug_read: copy");
    return RT_OK;
}

// out == NULL: enqueue the launch and return without waiting (timing).
extern "C" int rt_batch_retrack(rt_batch* b, const double* init_T, rt_pair_result* out) {
    if (!b || b->n < 1) return set_error(RT_ERR_STATE, "rt_batch_retrack: no prepared batch");
    if (hipSetDevice(b->device) != hipSuccess) return set_error(RT_ERR_HIP, "hipSetDevice");
    if (init_T) {
        std::vector<float> trel(16 * (size_t)b->n);
        for (size_t i = 0; i < trel.size(); i++) trel[i] = (float)init_T[i];
        if (!b->st_trel.upload(b->trel_in, trel.data(), 64 * (size_t)b->n, b->stream))
            return set_error(RT_ERR_HIP, "pose upload");
    }
    return run_track(b, out);
}

extern "C" int rt_batch_results(rt_batch* b, rt_pair_result* out) {
    if (!b || b->n < 1 || !out) return set_error(RT_ERR_STATE, "rt_batch_results: no prepared batch or no output");
    if (hipSetDevice(b->device) != hipSuccess) return set_error(RT_ERR_HIP, "hipSetDevice");
    return fetch_results(b, out);
}

// rcp_depth(-u) for u = 1 .. n, the reciprocal kb_bins takes of each 16-bit depth
__global__ __launch_bounds__(256) void k_depth_reciprocals(float* out, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = rcp_depth(-(float)(i + 1));
}
extern "C" int rt_debug_depth_reciprocals(int device, float* host, int n) {
    if (!host || n < 1 || n > 65535) return set_error(RT_ERR_ARG, "rt_debug_depth_reciprocals: n outside [1, 65535]");
    if (hipSetDevice(device) != hipSuccess) return set_error(RT_ERR_HIP, "hipSetDevice");
    float* d = nullptr;
    if (hipMalloc(&d, 4 * (size_t)n) != hipSuccess) return set_error(RT_ERR_HIP, "rt_debug_depth_reciprocals: alloc");
    k_depth_reciprocals<<<(n + 255) / 256, 256>>>(d, n);
    const bool ok = hipGetLastError() == hipSuccess && hipMemcpy(host, d, 4 * (size_t)n, hipMemcpyDeviceToHost) == hipSuccess;
    hipFree(d);
    return ok ? RT_OK : set_error(RT_ERR_HIP, "rt_debug_depth_reciprocals: launch or copy");
}

// ------------------------------------------------------------------ no-alloc batch tracking
// rt_batch_track for callers that cannot allocate device memory: the pair
// arenas and tables live in one fixed device array (`scratch`, held by every
// GPU the library runs on until the process ends) instead of a batch's own
// allocations; the caller passes frames, calibration and results as for
// rt_batch_track. Limits: RT_NOALLOC_MAX_PAIRS pairs of at most
// RT_NOALLOC_MAX_WIDTH x RT_NOALLOC_MAX_HEIGHT and RT_NOALLOC_MAX_POINTS points.
namespace rt {
constexpr size_t ca256(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr size_t NA_PAIR_BYTES = 2 * ca256((size_t)RT_NOALLOC_MAX_WIDTH * RT_NOALLOC_MAX_HEIGHT * 4 / 3 + 64) +
                                 ca256((size_t)RT_NOALLOC_MAX_WIDTH * RT_NOALLOC_MAX_HEIGHT) +
                                 ca256((size_t)RT_NOALLOC_MAX_WIDTH * ((RT_NOALLOC_MAX_HEIGHT + 7) & ~7) * 2) + ca256(256 * 4) +
                                 ca256((size_t)RT_NOALLOC_MAX_POINTS * 4) + ca256((size_t)RT_NOALLOC_MAX_POINTS * 16) +
                                 4 * ca256((size_t)RT_NOALLOC_MAX_POINTS * 28);
// table region: descriptors, work order, queue, calibration blocks, poses, rms, status
constexpr size_t NA_DESC = 0, NA_ORDER = 4096, NA_QUEUE = 4352, NA_CALIB = 4608,
                 NA_TREL_IN = NA_CALIB + ca256((size_t)RT_NOALLOC_MAX_PAIRS * 96 * 4),
                 NA_TREL_OUT = NA_TREL_IN + ca256((size_t)RT_NOALLOC_MAX_PAIRS * 64),
                 NA_RMS = NA_TREL_OUT + ca256((size_t)RT_NOALLOC_MAX_PAIRS * 64),
                 NA_STATUS = NA_RMS + ca256((size_t)RT_NOALLOC_MAX_PAIRS * 4),
                 NA_TABLES = ca256(NA_STATUS + (size_t)RT_NOALLOC_MAX_PAIRS * 4);
static_assert(sizeof(PairDesc) * RT_NOALLOC_MAX_PAIRS <= NA_ORDER, "descriptor table");
static_assert(NA_ORDER + 4 * RT_NOALLOC_MAX_PAIRS <= NA_QUEUE, "order table");
constexpr size_t NA_SCRATCH_BYTES = NA_TABLES + (size_t)RT_NOALLOC_MAX_PAIRS * NA_PAIR_BYTES;
static_assert(NA_SCRATCH_BYTES == RT_NOALLOC_SCRATCH_BYTES, "the header states the scratch size");
}  // namespace rt

// Host code never uses this name as an address: the per-device address comes
// from hipGetSymbolAddress (noalloc_scratch below).
static __device__ __align__(256) unsigned char scratch[rt::NA_SCRATCH_BYTES];

namespace {
std::mutex g_na_mutex;
void* g_na_ptr[64] = {};
// The device address of `scratch` on the current device, looked up once per
// device (with that device current).
int noalloc_scratch(int device, char** out) {
    std::lock_guard<std::mutex> lk(g_na_mutex);
    if (!g_na_ptr[device]) {
        void* ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
            (void)hipGetLastError();
            return set_error(RT_ERR_HIP, "rt_batch_track_noalloc: hipGetSymbolAddress of the scratch array failed");
        }
        g_na_ptr[device] = ptr;
    }
    *out = (char*)g_na_ptr[device];
    return RT_OK;
}
// The tracking kernel variant of the no-alloc path (no autotuning): 1024
// threads with residuals and the point cache in LDS where it covers the
// points, else 1024 threads with residuals in registers.
int noalloc_variant(int maxN) {
    int fallback = -1;
    for (int v = 0; v < NUM_VARIANTS; v++) {
        const rt::Variant& k = rt::kVariants[v];
        if (k.bs != 1024 || k.mode != RT_PHOTOMETRIC) continue;
        if (k.rz && k.pc && k.ppt == 4 && lds_variant_ready(k, maxN)) return v;
        if (!k.rz && !k.pc && k.bs * k.ppt >= maxN && fallback < 0) fallback = v;
    }
    return fallback;
}
}  // namespace

extern "C" int rt_batch_track_noalloc(int device, int n, const rt_pair* pairs, const rt_params* params, void* stream,
                                      rt_pair_result* out) {
    if (!pairs || !out || n < 1) return set_error(RT_ERR_ARG, "rt_batch_track_noalloc: bad argument");
    if (n > RT_NOALLOC_MAX_PAIRS)
        return set_error(RT_ERR_ARG, "rt_batch_track_noalloc: more pairs than RT_NOALLOC_MAX_PAIRS");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev || device >= 64)
        return set_error(RT_ERR_ARG, "rt_batch_track_noalloc: no such device");
    if (hipSetDevice(device) != hipSuccess) return set_error(RT_ERR_HIP, "hipSetDevice");
    rt_params prm;
    if (params) prm = *params; else rt_default_params(&prm);
    if (prm.n_select < 256 || prm.n_select % 256 || prm.n_select > RT_NOALLOC_MAX_POINTS)
        return set_error(RT_ERR_ARG, "rt_batch_track_noalloc: n_select must be a multiple of 256 up to RT_NOALLOC_MAX_POINTS");
    if (prm.mad_bins != 64 && prm.mad_bins != 256) return set_error(RT_ERR_ARG, "mad_bins must be 64 or 256");
    if (prm.mode != RT_PHOTOMETRIC) return set_error(RT_ERR_UNSUPPORTED, "rt_batch_track_noalloc: photometric mode only");
    for (int i = 0; i < n; i++)
        if (pairs[i].width > RT_NOALLOC_MAX_WIDTH || pairs[i].height > RT_NOALLOC_MAX_HEIGHT)
            return set_error(RT_ERR_ARG, "rt_batch_track_noalloc: pair larger than RT_NOALLOC_MAX_WIDTH x RT_NOALLOC_MAX_HEIGHT");
    if (pair_arena_bytes(RT_NOALLOC_MAX_WIDTH, RT_NOALLOC_MAX_HEIGHT, RT_NOALLOC_MAX_POINTS) != NA_PAIR_BYTES)
        return set_error(RT_ERR_STATE, "rt_batch_track_noalloc: arena layout mismatch");
    char* base = nullptr;
    int rc = noalloc_scratch(device, &base);
    if (rc) return rc;
    const BatchMem m{base + NA_TABLES, NA_PAIR_BYTES, RT_NOALLOC_MAX_WIDTH, RT_NOALLOC_MAX_HEIGHT,
                     (size_t)RT_NOALLOC_MAX_POINTS, false, (float*)(base + NA_CALIB), (float*)(base + NA_TREL_IN),
                     (float*)(base + NA_TREL_OUT), (float*)(base + NA_RMS), (int*)(base + NA_STATUS)};
    // host tables in this frame; the stream is synchronised before return, so
    // the queued copies have read them by then
    PairDesc desc[RT_NOALLOC_MAX_PAIRS];
    float calib[RT_NOALLOC_MAX_PAIRS * 96], trel[RT_NOALLOC_MAX_PAIRS * 16];
    int order[RT_NOALLOC_MAX_PAIRS];
    int maxW = 0, maxH = 0, maxN = 0;
    if ((rc = build_batch(prm, m, n, pairs, desc, calib, trel, order, maxW, maxH, maxN))) return rc;
    const int variant = noalloc_variant(maxN);
    if (variant < 0) return set_error(RT_ERR_UNSUPPORTED, "rt_batch_track_noalloc: no kernel variant covers the point count");
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus < 1) cus = 256;
    hipStream_t s = (hipStream_t)stream;
    if (hipMemcpyAsync(base + NA_DESC, desc, sizeof(PairDesc) * n, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(base + NA_ORDER, order, 4 * (size_t)n, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(base + NA_CALIB, calib, 96 * 4 * (size_t)n, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(base + NA_TREL_IN, trel, 64 * (size_t)n, hipMemcpyHostToDevice, s) != hipSuccess) {
        (void)hipStreamSynchronize(s);
        return set_error(RT_ERR_HIP, "rt_batch_track_noalloc: table upload");
    }
    const BatchView v{(const PairDesc*)(base + NA_DESC), (const int*)(base + NA_ORDER), (int*)(base + NA_QUEUE),
                      n, maxW, maxH, maxN, prm.solver, prm.robust, prm.mad_bins, prm.mode, prm.lambda};
    (void)hipGetLastError();
    launch_prep(v, s);
    launch_track(v, variant, std::min(n, cus), s);
    hipError_t le = hipGetLastError();
    float tr[RT_NOALLOC_MAX_PAIRS * 16], rms[RT_NOALLOC_MAX_PAIRS];
    int st[RT_NOALLOC_MAX_PAIRS];
    hipError_t ce = hipMemcpyAsync(tr, m.trel_out, 64 * (size_t)n, hipMemcpyDeviceToHost, s);
    if (ce == hipSuccess) ce = hipMemcpyAsync(rms, m.rms, 4 * (size_t)n, hipMemcpyDeviceToHost, s);
    if (ce == hipSuccess) ce = hipMemcpyAsync(st, m.status, 4 * (size_t)n, hipMemcpyDeviceToHost, s);
    const hipError_t se = hipStreamSynchronize(s);
    if (le != hipSuccess) return set_error(RT_ERR_HIP, "rt_batch_track_noalloc: launch failed");
    if (ce != hipSuccess || se != hipSuccess) return set_error(RT_ERR_HIP, "rt_batch_track_noalloc: result copy");
    for (int i = 0; i < n; i++) {
        double Td[16], Ti[16];
        for (int k = 0; k < 16; k++) Td[k] = tr[16 * i + k];
        invert_rt4(Td, Ti);
        memcpy(out[i].Trel, Td, sizeof(Td));
        memcpy(out[i].T, Ti, sizeof(Ti));
        out[i].rms = rms[i];
        out[i].status = st[i];
    }
    return RT_OK;
}

// ------------------------------------------------------------------ layout inspection
// The descriptors rt_batch_prepare would build for n pairs in a batch of
// max_width x max_height, as offsets from the batch's arena (host code only:
// no device is touched), for bounds checks of the launches on the CPU. Per pair
// 40 int64 values: W, H, L, N, simple, then the byte offsets of gray[0][1..3],
// gray[1][1..3], bins, ghist, sel, pts, jac[0..3] from the arena, the pair's
// arena slot start and the slot size, then (geometric modes, else -1) the
// offsets of rpos, rnrm, cpos of levels 0..3, isel, ir2, icnt, ivar; then the
// work order.
extern "C" int rt_batch_layout(const rt_params* params, int max_w, int max_h, int n, const rt_pair* pairs,
                               int64_t* fields, int* order, int maxes[3]) {
    if (!pairs || !fields || !order || !maxes || n < 1) return set_error(RT_ERR_ARG, "rt_batch_layout: bad argument");
    rt_params prm;
    if (params) prm = *params; else rt_default_params(&prm);
    // a nominal arena address: only offsets from it are reported
    char* const arena = reinterpret_cast<char*>((uintptr_t)1 << 40);
    const bool icp = prm.mode != RT_PHOTOMETRIC;
    const size_t pb = pair_arena_bytes(max_w, max_h, (size_t)prm.n_select, icp);
    const BatchMem m{arena, pb, max_w, max_h, (size_t)prm.n_select, icp, nullptr, nullptr, nullptr, nullptr, nullptr};
    std::vector<PairDesc> desc(n);
    std::vector<float> calib(96 * (size_t)n), trel(16 * (size_t)n);
    int rc = build_batch(prm, m, n, pairs, desc.data(), calib.data(), trel.data(), order, maxes[0], maxes[1], maxes[2]);
    if (rc) return rc;
    auto off = [&](const void* p) { return p ? (int64_t)((const char*)p - arena) : (int64_t)-1; };
    for (int i = 0; i < n; i++) {
        const PairDesc& d = desc[i];
        int64_t* f = fields + (size_t)i * 40;
        f[0] = d.W; f[1] = d.H; f[2] = d.L; f[3] = d.N; f[4] = d.simple;
        for (int l = 1; l < 4; l++) { f[4 + l] = off(d.gray[0][l]); f[7 + l] = off(d.gray[1][l]); }
        f[11] = off(d.bins); f[12] = off(d.ghist); f[13] = off(d.sel); f[14] = off(d.pts);
        for (int l = 0; l < 4; l++) f[15 + l] = off(d.jac[l]);
        f[19] = (int64_t)(pb * (size_t)i); f[20] = (int64_t)pb;
        for (int l = 0; l < 4; l++) {
            f[21 + 3 * l] = icp ? off(d.rpos[l]) : -1;
            f[22 + 3 * l] = icp ? off(d.rnrm[l]) : -1;
            f[23 + 3 * l] = icp ? off(d.cpos[l]) : -1;
        }
        f[33] = icp ? off(d.isel) : -1; f[34] = icp ? off(d.ir2) : -1;
        f[35] = icp ? off(d.icnt) : -1; f[36] = icp ? off(d.ivar) : -1;
        f[37] = off(d.zblk);
        f[38] = f[39] = 0;
    }
    return RT_OK;
}
