// rt_nn.h -- learned denoiser: sizes, parameter layout and launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rtk {

constexpr int kNnFeat = 16;     // input feature channels (10 used)
constexpr int kNnHidden = 32;   // hidden channels
constexpr int kNnLayers = 4;    // 3x3 convolutions, dilation 1, 2, 4, 8
constexpr int kNnHeadParams = 3 * kNnHidden + 3;
constexpr int kNnMaxImages = 16;
constexpr int kNnGradPixels = 2048;  // pixels per weight-gradient block

// parameter vector: per layer W[co][tap][ci] then b[co]; then head W[c][k], b[c]
__host__ __device__ constexpr int nn_layer_params(int l) {
    return kNnHidden * 9 * (l == 0 ? kNnFeat : kNnHidden) + kNnHidden;
}
__host__ __device__ constexpr int nn_layer_offset(int l) {
    int o = 0;
    for (int i = 0; i < l; ++i) o += nn_layer_params(i);
    return o;
}
__host__ __device__ constexpr int nn_head_offset() { return nn_layer_offset(kNnLayers); }
__host__ __device__ constexpr int nn_param_count() { return nn_head_offset() + kNnHeadParams; }

// images of one batch, back to back
struct NnImages {
    int32_t count;
    int32_t offset[kNnMaxImages + 1];
    int32_t width[kNnMaxImages];
};

// device weights in the kernels' layouts
struct NnPacked {
    uint16_t* fwd[kNnLayers];  // bf16 [32][KP]
    uint16_t* bwd[kNnLayers];  // bf16 [CIN][9 * 32] (layers >= 1)
    float* bias[kNnLayers];
    float* head;               // FP32 W[3][32], b[3]
};

struct NnScratch {
    uint16_t* feat;               // bf16 [P][16]
    uint16_t* act[kNnLayers];     // bf16 [P][32] post-ReLU
    uint16_t* dact[2];            // bf16 [P][32] pre-activation gradients (ping-pong)
    float* head_partial;          // [blocks of 256 px][kNnHeadParams + 1]
    float* grad_partial;          // [blocks of kNnGradPixels px][32][19 * 16]
};

hipError_t nn_pack(const float* params, const NnPacked& pk, hipStream_t st);
// features, four convolutions and the head; with `target` also the loss
// terms and the head-level gradients needed by nn_backward
hipError_t nn_forward(const NnImages& im, int P, const NnPacked& pk, const NnScratch& w,
    const float4* beauty, const float4* normal, const float4* albedo, const float* depth,
    const float4* target, float4* out, int conv_grid, hipStream_t st);
// parameter gradient (FP32, parameter layout) and loss of the last forward
hipError_t nn_backward(const NnImages& im, int P, const NnPacked& pk, const NnScratch& w,
    float* grad, float* loss, int conv_grid, hipStream_t st);
hipError_t nn_adam(int n, float* params, const float* grad, float* m, float* v, float lr, int step,
    hipStream_t st);

}  // namespace rtk
