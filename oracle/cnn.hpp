// CPU oracle of the learned pose proposal (TEST INFRASTRUCTURE ONLY).
//
// The network has no counterpart in the reference: it is the small CNN that
// warm-starts Gauss-Newton at large motions. Its forward and backward passes
// are written here once, in FP64, as the checker of the HIP kernels; the
// backward pass is itself pinned by finite differences of the FP64 forward.
//
// Network: NHWC input of C0 channels at H0 x W0; six 3x3 convolutions with
// stride 2 and zero padding 1 (out = ceil(in / 2)), each followed by ReLU; a
// global average pool; a fully connected layer to 6 outputs.
// Flat parameter layout (FP32): for each conv layer l: W_l[Cout][3][3][Cin]
// then b_l[Cout]; then Wfc[6][C6] and bfc[6].
// Loss: sum_n w_n * sum_j (out_nj - y_nj)^2 / max(1, sum_n w_n).
//
// round = true emulates the device's storage formats: conv weights, inputs,
// activations and the back-propagated conv output gradients are rounded to
// bfloat16 (round to nearest even) where the device stores them as bfloat16;
// every sum is exact-ish FP64 here (the device accumulates in FP32).
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace oracle {
namespace cnn {

inline float bf16_round(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) {  // inf / nan: truncate, keep nan quiet
        if (u & 0x007fffffu) u |= 0x00400000u;
        u &= 0xffff0000u;
    } else {
        u += 0x7fffu + ((u >> 16) & 1u);
        u &= 0xffff0000u;
    }
    std::memcpy(&x, &u, 4);
    return x;
}

struct Net {
    int ch[7], H[7], W[7];
    size_t w_off[6], b_off[6], fc_w, fc_b, nparams;
    Net(const int* chans, int H0, int W0) {
        for (int l = 0; l < 7; l++) ch[l] = chans[l];
        H[0] = H0; W[0] = W0;
        for (int l = 1; l < 7; l++) { H[l] = (H[l - 1] + 1) / 2; W[l] = (W[l - 1] + 1) / 2; }
        size_t o = 0;
        for (int l = 0; l < 6; l++) {
            w_off[l] = o; o += (size_t)ch[l + 1] * 9 * ch[l];
            b_off[l] = o; o += (size_t)ch[l + 1];
        }
        fc_w = o; o += 6 * (size_t)ch[6];
        fc_b = o; o += 6;
        nparams = o;
    }
};

// Loss of a batch; out (B*6) receives the network outputs; grad (nparams),
// when given, receives d loss / d params.
inline double forward_backward(const Net& net, int B, const float* params, const float* x, const float* y,
                               const float* wn, bool round, double* out, double* grad) {
    auto rd = [&](float v) -> double { return round ? (double)bf16_round(v) : (double)v; };
    // activations a[0..6], each B*H*W*C
    std::vector<std::vector<double>> a(7);
    size_t n0 = (size_t)B * net.H[0] * net.W[0] * net.ch[0];
    a[0].resize(n0);
    for (size_t i = 0; i < n0; i++) a[0][i] = rd(x[i]);
    std::vector<std::vector<double>> Wr(6);
    for (int l = 0; l < 6; l++) {
        size_t nw = (size_t)net.ch[l + 1] * 9 * net.ch[l];
        Wr[l].resize(nw);
        for (size_t i = 0; i < nw; i++) Wr[l][i] = rd(params[net.w_off[l] + i]);
    }
    for (int l = 0; l < 6; l++) {
        int Ci = net.ch[l], Co = net.ch[l + 1], Hi = net.H[l], Wi = net.W[l], Ho = net.H[l + 1], Wo = net.W[l + 1];
        a[l + 1].assign((size_t)B * Ho * Wo * Co, 0.0);
        const float* bias = params + net.b_off[l];
        for (int n = 0; n < B; n++)
            for (int yo = 0; yo < Ho; yo++)
                for (int xo = 0; xo < Wo; xo++)
                    for (int co = 0; co < Co; co++) {
                        double s = 0.0;
                        for (int ky = 0; ky < 3; ky++) {
                            int yi = 2 * yo - 1 + ky;
                            if (yi < 0 || yi >= Hi) continue;
                            for (int kx = 0; kx < 3; kx++) {
                                int xi = 2 * xo - 1 + kx;
                                if (xi < 0 || xi >= Wi) continue;
                                const double* ap = &a[l][(((size_t)n * Hi + yi) * Wi + xi) * Ci];
                                const double* wp = &Wr[l][(((size_t)co * 3 + ky) * 3 + kx) * Ci];
                                for (int ci = 0; ci < Ci; ci++) s += ap[ci] * wp[ci];
                            }
                        }
                        double z = s + (double)bias[co];
                        double r = z > 0.0 ? z : 0.0;
                        a[l + 1][(((size_t)n * Ho + yo) * Wo + xo) * Co + co] = round ? (double)bf16_round((float)r) : r;
                    }
    }
    int C6 = net.ch[6], P6 = net.H[6] * net.W[6];
    std::vector<double> h((size_t)B * C6, 0.0);
    for (int n = 0; n < B; n++)
        for (int p = 0; p < P6; p++)
            for (int c = 0; c < C6; c++) h[(size_t)n * C6 + c] += a[6][((size_t)n * P6 + p) * C6 + c];
    for (auto& v : h) v /= P6;
    const float* Wfc = params + net.fc_w;
    const float* bfc = params + net.fc_b;
    double wsum = 0.0;
    for (int n = 0; n < B; n++) wsum += wn[n];
    double norm = wsum > 1.0 ? wsum : 1.0;
    double loss = 0.0;
    std::vector<double> dout((size_t)B * 6);
    for (int n = 0; n < B; n++)
        for (int j = 0; j < 6; j++) {
            double o = bfc[j];
            for (int c = 0; c < C6; c++) o += (double)Wfc[j * C6 + c] * h[(size_t)n * C6 + c];
            out[n * 6 + j] = o;
            double e = o - y[n * 6 + j];
            loss += wn[n] * e * e / norm;
            dout[n * 6 + j] = 2.0 * wn[n] * e / norm;
        }
    if (!grad) return loss;
    std::memset(grad, 0, sizeof(double) * net.nparams);
    std::vector<double> dh((size_t)B * C6, 0.0);
    for (int n = 0; n < B; n++)
        for (int j = 0; j < 6; j++) {
            double d = dout[n * 6 + j];
            grad[net.fc_b + j] += d;
            for (int c = 0; c < C6; c++) {
                grad[net.fc_w + j * C6 + c] += d * h[(size_t)n * C6 + c];
                dh[(size_t)n * C6 + c] += (double)Wfc[j * C6 + c] * d;
            }
        }
    // gradient w.r.t. the output of layer 6 (post-ReLU activations)
    std::vector<double> da((size_t)B * P6 * C6);
    for (int n = 0; n < B; n++)
        for (int p = 0; p < P6; p++)
            for (int c = 0; c < C6; c++) da[((size_t)n * P6 + p) * C6 + c] = dh[(size_t)n * C6 + c] / P6;
    for (int l = 5; l >= 0; l--) {
        int Ci = net.ch[l], Co = net.ch[l + 1], Hi = net.H[l], Wi = net.W[l], Ho = net.H[l + 1], Wo = net.W[l + 1];
        // through the ReLU: dz = da where the activation is positive
        std::vector<double> dz(da.size());
        for (size_t i = 0; i < da.size(); i++) {
            double v = a[l + 1][i] > 0.0 ? da[i] : 0.0;
            dz[i] = round ? (double)bf16_round((float)v) : v;
        }
        std::vector<double> dprev;
        if (l > 0) dprev.assign((size_t)B * Hi * Wi * Ci, 0.0);
        double* gW = grad + net.w_off[l];
        double* gb = grad + net.b_off[l];
        for (int n = 0; n < B; n++)
            for (int yo = 0; yo < Ho; yo++)
                for (int xo = 0; xo < Wo; xo++)
                    for (int co = 0; co < Co; co++) {
                        double d = dz[(((size_t)n * Ho + yo) * Wo + xo) * Co + co];
                        if (d == 0.0) continue;
                        gb[co] += d;
                        for (int ky = 0; ky < 3; ky++) {
                            int yi = 2 * yo - 1 + ky;
                            if (yi < 0 || yi >= Hi) continue;
                            for (int kx = 0; kx < 3; kx++) {
                                int xi = 2 * xo - 1 + kx;
                                if (xi < 0 || xi >= Wi) continue;
                                size_t ao = (((size_t)n * Hi + yi) * Wi + xi) * Ci;
                                size_t wo = (((size_t)co * 3 + ky) * 3 + kx) * Ci;
                                for (int ci = 0; ci < Ci; ci++) {
                                    gW[wo + ci] += d * a[l][ao + ci];
                                    if (l > 0) dprev[ao + ci] += d * Wr[l][wo + ci];
                                }
                            }
                        }
                    }
        da.swap(dprev);
    }
    return loss;
}

}  // namespace cnn
}  // namespace oracle
