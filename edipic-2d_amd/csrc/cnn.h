// Learned warm start network: buffers and launchers (cnn.hip).
#pragma once

#include "kernels.h"

namespace edgpu {

constexpr int kCnnLayers = 6;
constexpr int kCnnWgradBlocks = 256;

// input/target scales (root mean squares over the free nodes of the recorded pairs): residual,
// potential increment phi^{n-1} - phi^{n-2}, right-hand side, target phi^n - phi^{n-1}
struct CnnScales {
    double s_r, s_i, s_b, s_d;
};

struct CnnNet {
    int64_t nparam = 0;
    float *master = nullptr, *grad = nullptr, *m1 = nullptr, *m2 = nullptr;  // FP32, one vector
    uint16_t *wbf = nullptr;                                                 // BF16 copy, same layout
    uint8_t *frozen = nullptr;
    int64_t off_w[kCnnLayers], off_b[kCnnLayers];
    uint16_t *wb[kCnnLayers];  // BF16 weights of layer l: [tap][cout][cin]
    float *b[kCnnLayers];      // FP32 bias
    float *gw[kCnnLayers], *gb[kCnnLayers];
    int t = 0;                 // Adam step
};

struct CnnActs {
    uint16_t *feat = nullptr;            // 16 channels
    uint16_t *act[kCnnLayers - 1] = {};  // 32 channels each
    float *pred = nullptr;               // 1 channel
    uint16_t *dout16 = nullptr;          // dL/d(last layer output), 16 channels
    uint16_t *g[2] = {};                 // 32-channel gradient ping-pong
    uint16_t *wflip = nullptr;           // 9*32*32
    float *partial = nullptr;            // kCnnWgradBlocks * (9*32*32 + 32)
    float *gsum = nullptr;               // 9*32*32 + 32
};

void cnn_layer_shape(int l, int &cin, int &cout);
void cnn_forward(const CnnNet &net, CnnActs &a, int H, int W, hipStream_t s);
hipError_t cnn_backward(CnnNet &net, CnnActs &a, int H, int W, hipStream_t s);
// MSE of a.pred against target over the free nodes (feature channel 3); fills a.dout16
hipError_t cnn_loss(CnnActs &a, const float *target, int64_t n, int64_t nfree, double *dloss, double *loss,
                    hipStream_t s);
void cnn_adam(CnnNet &net, float lr, hipStream_t s);
void cnn_sync_bf16(CnnNet &net, hipStream_t s);
// raw FP32 features (residual, increment, rhs, free mask; 4 per node) of the solve for b
// starting from phi1 (phi2 the potential before it)
void launch_cnn_features_raw(const double *b, const double *phi1, const double *phi2, const uint8_t *mask,
                             const Geo &g, float *raw4, hipStream_t s);
// raw FP32 target phi - phi1 on the free nodes
void launch_cnn_target_raw(const double *phi, const double *phi1, const uint8_t *mask, float *t, int64_t n,
                           hipStream_t s);
// over the free nodes of n nodes: acc[0..2] += sums of squares of the residual, increment
// and right-hand-side channels, acc[3] += of the target, acc[4] += count (FP64)
void launch_cnn_sumsq(const float *raw4, const float *t, int64_t n, double *acc, hipStream_t s);
// normalised BF16 network input (16 channels) and FP32 target from raw ones
void launch_cnn_pack(const float *raw4, const CnnScales &sc, uint16_t *feat, int64_t n, hipStream_t s);
void launch_cnn_scale(const float *t, float inv, float *out, int64_t n, hipStream_t s);
void launch_cnn_apply(const float *pred, const double *phi1, const uint8_t *mask, double s_d, double *x, int64_t n,
                      hipStream_t s);
void launch_mfma_probe(const float *A, const float *B, float *D, hipStream_t s);

}  // namespace edgpu
