// Internal interface of the pose-proposal network (cnn.hip) for the batch
// tracker (rt_batch.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rt_cnn.h"

namespace rt {

typedef uint16_t bf16;
// float to bfloat16, round to nearest even; NaN stays quiet NaN
__device__ __forceinline__ bf16 f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) {
        if (u & 0x7fffffu) u |= 0x400000u;
        return (bf16)(u >> 16);
    }
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16)(u >> 16);
}

// The network's input features of one frame over a cell of npx pixels: the
// mean gray / 255 - 0.5 from the cell's integer gray sum gs, and the mean
// inverse depth - 1 over the nv valid depths (ids: the sum of their
// inv_depth_m), or -1 without valid depth. Every builder of the inputs sums
// ids per column over the cell's rows in order and combines the column sums
// pairwise (adjacent columns, then adjacent pairs, ...), so that all give
// the same bits: the input builder (k_inputs) and the batch preparation's
// fused pass on cells of 8 x 8 pixels.
__device__ __forceinline__ float inv_depth_m(unsigned d) { return 1000.0f * __builtin_amdgcn_rcpf((float)d); }
__device__ __forceinline__ void cell_features(unsigned gs, float ids, unsigned nv, int npx, bf16* o) {
    o[0] = f2bf((float)gs / (255.0f * (float)npx) - 0.5f);
    o[1] = f2bf(nv ? ids / (float)nv - 1.0f : -1.0f);
}

// cell_features of the 8 x 8 cell with top-left pixel (x0, y0) of one frame:
// one 8-byte gray and one 16-byte depth load per row (x0 and W multiples of
// 8, the gray 8-byte and the depth 16-byte aligned); column sums over the
// rows in order, then ((c0 + c1) + (c2 + c3)) + ((c4 + c5) + (c6 + c7)).
__device__ __forceinline__ void cell8_features(const uint8_t* g8, const uint16_t* d16, int W, int x0, int y0, bf16* o) {
    unsigned gs = 0, nv = 0;
    float col[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    uint2 gq[8];
    uint4 dq[8];
#pragma unroll
    for (int yy = 0; yy < 8; yy++) {  // all 16 loads in flight together
        const size_t row = (size_t)(y0 + yy) * W + x0;
        gq[yy] = *(const uint2*)(g8 + row);
        dq[yy] = *(const uint4*)(d16 + row);
    }
#pragma unroll
    for (int yy = 0; yy < 8; yy++) {
        const unsigned gw[2] = {gq[yy].x, gq[yy].y};
        const unsigned dw[4] = {dq[yy].x, dq[yy].y, dq[yy].z, dq[yy].w};
#pragma unroll
        for (int xx = 0; xx < 8; xx++) {
            gs += (gw[xx >> 2] >> (8 * (xx & 3))) & 0xffu;
            const unsigned d = (dw[xx >> 1] >> (16 * (xx & 1))) & 0xffffu;
            if (d) { col[xx] += inv_depth_m(d); nv++; }
        }
    }
    const float ids = ((col[0] + col[1]) + (col[2] + col[3])) + ((col[4] + col[5]) + (col[6] + col[7]));
    cell_features(gs, ids, nv, 64, o);
}

// The raw inputs of one frame pair as the network's input builder reads them.
struct FrameSrc {
    int W, H;
    const uint8_t* g8[2];    // reference, current gray
    const uint16_t* d16[2];  // reference, current depth (mm)
};

// Proposals for n pairs on stream s: builds inputs from src (host array),
// runs the forward pass and writes each pair's initial Trel (16 floats,
// row-major) to trel_dev + 16 i. x_ready: the inputs already built on the
// device (cells of 8 x 8 pixels, the layout k_inputs writes), or null.
int cnn_propose(rt_cnn* net, int n, const FrameSrc* src, float* trel_dev, hipStream_t s, const bf16* x_ready = nullptr);
// whether the network's input cells are 8 x 8 pixels of a W x H frame
bool cnn_cells8(const rt_cnn* net, int W, int H);

}  // namespace rt
