// Internal interface of the pose-proposal network (cnn.hip) for the batch
// tracker (rt_batch.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rt_cnn.h"

namespace rt {

// The raw inputs of one frame pair as the network's input builder reads them.
struct FrameSrc {
    int W, H;
    const uint8_t* g8[2];    // reference, current gray
    const uint16_t* d16[2];  // reference, current depth (mm)
};

// Proposals for n pairs on stream s: builds inputs from src (host array),
// runs the forward pass and writes each pair's initial Trel (16 floats,
// row-major) to trel_dev + 16 i.
int cnn_propose(rt_cnn* net, int n, const FrameSrc* src, float* trel_dev, hipStream_t s);

}  // namespace rt
