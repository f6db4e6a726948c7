// rt_nn.h -- learned denoiser: a 4-level convolutional encoder-decoder
// (widths 32 / 48 / 64 / 96, 3x3 convolutions, ReLU, skip connections),
// its parameter layout and launchers.
//
//   level 0 (full size):  e0a = relu(conv(f; 16->32)),  e0 = relu(conv(e0a; 32->32))
//   level 1 (1/2):        e1a = relu(conv(pool(e0); 32->48)), e1 = relu(conv(e1a; 48->48))
//   level 2 (1/4):        e2a = relu(conv(pool(e1); 48->64)), e2 = relu(conv(e2a; 64->64))
//   level 3 (1/8):        b0 = relu(conv(pool(e2); 64->96)), b1, b = relu(conv(.; 96->96)) x2
//   level 2:  d2a = relu(conv([e2, up(b)]; 160->64)),   d2 = relu(conv(d2a; 64->64))
//   level 1:  d1a = relu(conv([e1, up(d2)]; 112->48)),  d1 = relu(conv(d1a; 48->48))
//   level 0:  d0a = relu(conv([e0, up(d1)]; 80->32)),   d0 = relu(conv(d0a; 32->32))
//   out_log = log1p(beauty) + W_head d0 + b_head  (1x1, FP32)
//   loss = mean over pixels and RGB of w (out_log - log1p(t))^2, t = max(target, 0),
//          w = (1 + t)^2 / (t^2 + 1e-2) (relative squared error to first order)
// A level's image is ceil(w / 2) x ceil(h / 2) of the level above; pool is
// the mean of the (1, 2 or 4) pixels a level-(l+1) pixel covers; up(x) at a
// pixel is x at its parent.  Input features (16 channels, 10 used):
// log1p(beauty), albedo, normal * 2 - 1, log1p(depth) / 8.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rtk {

constexpr int kNnFeat = 16;
constexpr int kNnLevels = 4;
constexpr int kNnConvs = 15;
constexpr int kNnOut = 32;        // channels of d0 (the head's input)
constexpr int kNnHeadParams = 3 * kNnOut + 3;
constexpr int kNnMaxImages = 16;

// convolution l: c0 channels from tensor A at its level, c1 channels from
// tensor B one level coarser (upsampled), cout outputs
struct NnConv {
    int c0, c1, cout, level;
};
__host__ __device__ constexpr NnConv nn_conv(int l) {
    constexpr NnConv t[kNnConvs] = {{16, 0, 32, 0}, {32, 0, 32, 0}, {32, 0, 48, 1}, {48, 0, 48, 1},
        {48, 0, 64, 2}, {64, 0, 64, 2}, {64, 0, 96, 3}, {96, 0, 96, 3}, {96, 0, 96, 3},
        {64, 96, 64, 2}, {64, 0, 64, 2}, {48, 64, 48, 1}, {48, 0, 48, 1}, {32, 48, 32, 0},
        {32, 0, 32, 0}};
    return t[l];
}
__host__ __device__ constexpr int nn_cin(int l) { return nn_conv(l).c0 + nn_conv(l).c1; }
// parameter vector: per convolution W[cout][tap][cin] then b[cout]; then
// the head W[3][32], b[3]
__host__ __device__ constexpr int nn_layer_params(int l) {
    return nn_conv(l).cout * 9 * nn_cin(l) + nn_conv(l).cout;
}
__host__ __device__ constexpr int nn_layer_offset(int l) {
    int o = 0;
    for (int i = 0; i < l; ++i) o += nn_layer_params(i);
    return o;
}
__host__ __device__ constexpr int nn_head_offset() { return nn_layer_offset(kNnConvs); }
__host__ __device__ constexpr int nn_param_count() { return nn_head_offset() + kNnHeadParams; }
// K of a packed convolution GEMM over c input channels: k = tap * c + channel,
// the 9 c values padded once to a multiple of 32 (one MFMA K step); an
// 8-wide K group never straddles two taps since c is a multiple of 8
__host__ __device__ constexpr int nn_kp(int c) { return (9 * c + 31) / 32 * 32; }
__host__ __device__ constexpr int nn_kp_fwd(int l) { return nn_kp(nn_cin(l)); }
__host__ __device__ constexpr int nn_kp_bwd(int l) { return nn_kp(nn_conv(l).cout); }

// images of one batch, back to back
struct NnImages {
    int32_t count;
    int32_t offset[kNnMaxImages + 1];
    int32_t width[kNnMaxImages];
};

// the batch at every level (level 0 = the images)
struct NnLevels {
    NnImages lv[kNnLevels];
};

// device weights in the kernels' layouts
struct NnPacked {
    uint16_t* fwd[kNnConvs];   // bf16 [cout][kp_fwd], k = tap * cin + ci
    uint16_t* bwd[kNnConvs];   // bf16 [cin][kp_bwd], k = tap' * cout + co, W[co][8 - tap'][ci]
    float* bias[kNnConvs];
    float* head;               // FP32 W[3][32], b[3]
};

// activations (bf16, [pixels][channels]) and training buffers
struct NnScratch {
    uint16_t* feat;                  // level 0, 16 channels
    uint16_t* act[kNnConvs];         // output of convolution l
    uint16_t* pool[kNnLevels];       // pool[1..3]: pooled e0 / e1 / e2
    uint16_t* dz[kNnConvs];          // pre-activation gradient of convolution l
    uint16_t* r[kNnConvs];           // raw input gradient of convolution l (l >= 1)
    float* head_partial;             // [blocks of 256 px][kNnHeadParams + 1]
    float* grad_partial;             // weight-gradient partial sums
};

// rows per convolution tile at each level: 4 or 8 (more rows re-read fewer
// halo rows, fewer rows need less LDS per block); 0 chooses 8 when the
// level's widest image is at least 160 pixels, else 4
struct NnTiling {
    int32_t rows[kNnLevels];
};

NnLevels nn_levels(const NnImages& im);
// the tile rows convolutions at `level` (images `im`) run with
int nn_tile_rows(const NnTiling& t, const NnImages& im, int level);
hipError_t nn_pack(const float* params, const NnPacked& pk, hipStream_t st);
// features, the 15 convolutions and the head; with `target` also the loss
// terms and the head-level gradients needed by nn_backward
hipError_t nn_forward(const NnLevels& lv, const NnPacked& pk, const NnScratch& w,
    const float4* beauty, const float4* normal, const float4* albedo, const float* depth,
    const float4* target, float4* out, const NnTiling& tiling, hipStream_t st);
// parameter gradient (FP32, parameter layout) and loss of the last forward
hipError_t nn_backward(const NnLevels& lv, const NnPacked& pk, const NnScratch& w, float* grad,
    float* loss, const NnTiling& tiling, hipStream_t st);
hipError_t nn_adam(int n, float* params, const float* grad, float* m, float* v, float lr, int step,
    hipStream_t st);
// floats of weight-gradient partial sums needed for a batch
size_t nn_grad_partial_floats(const NnLevels& lv);

}  // namespace rtk
