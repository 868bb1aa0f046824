/*
 * rt_cnn.h -- learned pose proposal: a small CNN that warm-starts the
 * Gauss-Newton tracker at large motions.
 *
 * No counterpart in the reference (it tracks from the previous pose only,
 * phototracker.cpp:704-718); the plain tracker stays the fallback and the arbiter.
 * Convolutions run on the matrix cores (bfloat16 operands, FP32 accumulation);
 * the network is trained on the GPU by the forward and backward kernels behind
 * rt_cnn_train_step.
 *
 * Network: input C0 x in_height x in_width (NHWC bfloat16; the default 4
 * channels are gray and inverse depth of the reference and of the current frame
 * at 80x60), six 3x3 convolutions with stride 2, zero padding 1 and ReLU
 * (channels[1..6]), global average pooling, a fully connected layer to 6
 * outputs: the rotation vector of Trel in units of rot_scale_deg degrees and its
 * translation in units of trans_scale_mm millimetres.
 * Flat FP32 parameter layout: per conv layer W[Cout][3][3][Cin] then b[Cout];
 * then Wfc[6][C6] and bfc[6].
 * Loss of a batch: sum_n w_n |out_n - y_n|^2 / max(1, sum_n w_n).
 */
#ifndef RT_CNN_H
#define RT_CNN_H

#include <stddef.h>
#include <stdint.h>

#include "rt_tracker.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int in_width, in_height;   /* 80 x 60 */
    int channels[7];           /* 4, 32, 64, 128, 128, 256, 256; input 4 or a multiple of 4, the rest powers of two >= 16 */
    float rot_scale_deg;       /* 10 */
    float trans_scale_mm;      /* 150 */
} rt_cnn_config;

typedef struct rt_cnn rt_cnn;

void rt_cnn_default_config(rt_cnn_config* c);
/* max_batch: largest batch of any later call (inference and training). */
int rt_cnn_create(int device, const rt_cnn_config* c, int max_batch, rt_cnn** out);
void rt_cnn_destroy(rt_cnn* net);
size_t rt_cnn_num_params(const rt_cnn* net);
/* Host FP32 parameters in the flat layout; set also resets the optimiser state. */
int rt_cnn_set_params(rt_cnn* net, const float* params);
int rt_cnn_get_params(rt_cnn* net, float* params);
void* rt_cnn_stream(rt_cnn* net);

/* Network inputs of n frame pairs (device gray/depth pointers of rt_pair; the
 * calib and init_T fields are not used) into x_dev: n x in_height x in_width x
 * C0 bfloat16. Per cell of the (width/in_width) x (height/in_height) block:
 * mean gray / 255 - 0.5 and mean over valid depths of 1000/depth_mm - 1 (-1
 * where the block has no depth), reference frame then current frame. */
int rt_cnn_make_inputs(rt_cnn* net, int n, const rt_pair* pairs, uint16_t* x_dev);
/* Forward pass: out_dev receives n x 6 FP32 outputs (device). */
int rt_cnn_forward(rt_cnn* net, int n, const uint16_t* x_dev, float* out_dev);
/* Loss and gradients of a batch without updating (for checks): y_dev n x 6
 * targets, w_dev n sample weights (device); loss, out_host (n x 6) and
 * grad_host (num_params) are host outputs, each nullable. */
int rt_cnn_loss_grad(rt_cnn* net, int n, const uint16_t* x_dev, const float* y_dev, const float* w_dev, float* loss,
                     float* out_host, float* grad_host);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) at learning rate lr on the
 * batch; loss (host, nullable) receives the batch loss before the step. */
int rt_cnn_train_step(rt_cnn* net, int n, const uint16_t* x_dev, const float* y_dev, const float* w_dev, float lr,
                      float* loss);

/* Training augmentation by the camera's symmetries (principal point at the
 * image centre): per sample flags (host, n ints) bit 0 mirrors columns, bit 1
 * mirrors rows, bit 2 swaps the two frames; targets follow (y_inv_dev holds the
 * targets of the inverse poses). Writes x_out / y_out (device), enqueued on the
 * network's stream. */
int rt_cnn_augment(rt_cnn* net, int n, const uint16_t* x_dev, const float* y_dev, const float* y_inv_dev,
                   const int* flags, uint16_t* x_out, float* y_out);

/* Host conversions between poses and network outputs: targets from Trel
 * (row-major 4x4, ref -> cur point transform), and Trel from outputs. */
int rt_cnn_pose_to_target(const rt_cnn_config* c, int n, const double* Trel, float* y);
int rt_cnn_target_to_pose(const rt_cnn_config* c, int n, const float* y, double* Trel);

/* Batches: prepare the pairs (keyframes and pyramids) without tracking, then
 * let the network propose each pair's initial pose on the batch's stream;
 * rt_batch_retrack(b, NULL, out) then tracks from the proposals. */
int rt_batch_prepare(rt_batch* b, int n, const rt_pair* pairs);
int rt_batch_propose(rt_batch* b, rt_cnn* net);

#ifdef __cplusplus
}
#endif
#endif /* RT_CNN_H */
