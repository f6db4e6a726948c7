// rt_nn.hip -- the learned denoiser: BF16 MFMA inference and training
// kernels with FP32 accumulation (gfx950 v_mfma_f32_16x16x32_bf16).
//
// Network (per image, zero padding at borders; all images of a batch lie
// back to back):
//   f  (16 ch bf16) = log1p(beauty), albedo, normal, log1p(depth)/8, 0...
//   h0 = relu(conv3x3 dil 1 (f ; 16 -> 32))
//   h1 = relu(conv3x3 dil 2 (h0; 32 -> 32))
//   h2 = relu(conv3x3 dil 4 (h1; 32 -> 32))
//   h3 = relu(conv3x3 dil 8 (h2; 32 -> 32))       receptive field 31 x 31
//   out_log = log1p(beauty) + W_head h3 + b_head  (1x1, FP32)
//   loss = mean (out_log - log1p(target))^2
//
// Convolutions are implicit GEMMs: M = 16 pixels per MFMA tile, N = 32
// output channels (two 16-wide tiles), K = 9 taps x input channels in
// chunks of 32.  Each lane keeps its share of the layer's weights in
// registers for the whole launch and streams pixel tiles.  The input-gradient
// pass is the same kernel with transposed, tap-flipped weights and a ReLU
// mask epilogue; the weight-gradient pass is a GEMM with K = pixels, written
// as per-block partial sums reduced in a fixed order (deterministic).
#include <hip/hip_runtime.h>

#include "rt_bf16.h"
#include "rt_nn.h"

namespace rtk {
namespace {

struct Pix {
    int base, w, h, x, y;
};

__device__ __forceinline__ Pix locate(const NnImages& im, int p) {
    int i = 0;
    while (i + 1 < im.count && p >= im.offset[i + 1]) ++i;
    Pix r;
    r.base = im.offset[i];
    r.w = im.width[i];
    r.h = (im.offset[i + 1] - im.offset[i]) / r.w;
    const int local = p - r.base;
    r.y = local / r.w;
    r.x = local - r.y * r.w;
    return r;
}

__device__ __forceinline__ int tap_pixel(const Pix& q, int t, int d) {
    const int x = q.x + (t % 3 - 1) * d, y = q.y + (t / 3 - 1) * d;
    if (x < 0 || x >= q.w || y < 0 || y >= q.h) return -1;
    return q.base + y * q.w + x;
}

__global__ __launch_bounds__(256) void k_features(int P, const float4* beauty, const float4* normal,
    const float4* albedo, const float* depth, uint16_t* feat) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= P) return;
    const float4 b = beauty[p], n = normal[p], a = albedo[p];
    uint16_t f[kNnFeat];
    f[0] = to_bf16(log1pf(fmaxf(b.x, 0.0f)));
    f[1] = to_bf16(log1pf(fmaxf(b.y, 0.0f)));
    f[2] = to_bf16(log1pf(fmaxf(b.z, 0.0f)));
    f[3] = to_bf16(fminf(fmaxf(a.x, 0.0f), 1.0f));
    f[4] = to_bf16(fminf(fmaxf(a.y, 0.0f), 1.0f));
    f[5] = to_bf16(fminf(fmaxf(a.z, 0.0f), 1.0f));
    f[6] = to_bf16(n.x * 2.0f - 1.0f);
    f[7] = to_bf16(n.y * 2.0f - 1.0f);
    f[8] = to_bf16(n.z * 2.0f - 1.0f);
    f[9] = to_bf16(log1pf(fmaxf(depth[p], 0.0f)) * 0.125f);
    for (int c = 10; c < kNnFeat; ++c) f[c] = 0;
    uint4* dst = reinterpret_cast<uint4*>(feat + (size_t)p * kNnFeat);
    const uint32_t* w = reinterpret_cast<const uint32_t*>(f);
    dst[0] = make_uint4(w[0], w[1], w[2], w[3]);
    dst[1] = make_uint4(w[4], w[5], w[6], w[7]);
}

// Implicit-GEMM 3x3 convolution, 32 outputs.  MODE 0: out = relu(acc + bias);
// MODE 1: out = acc where mask > 0 (input gradient through a ReLU).
// W: bf16 [32][KP] with k = tap * CIN + ci (zero padded to KP).
template <int CIN, int MODE>
__global__ __launch_bounds__(256) void k_conv(NnImages im, int P, int dil, const uint16_t* in,
    const uint16_t* W, const float* bias, const uint16_t* mask, uint16_t* out) {
    constexpr int KC = (9 * CIN + 31) / 32;
    constexpr int KP = KC * 32;
    const int lane = threadIdx.x & 63;
    const int row = lane & 15;
    const int kq = lane >> 4;  // which 8-wide k group of the 32-wide chunk
    bf16x8 bw[KC][2];
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) {
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            bw[kc][nt] = *reinterpret_cast<const bf16x8*>(W + (size_t)(nt * 16 + row) * KP + kc * 32 + kq * 8);
        }
    }
    const int waves = gridDim.x * 4;
    const int tiles = (P + 15) / 16;
    for (int tile = blockIdx.x * 4 + (threadIdx.x >> 6); tile < tiles; tile += waves) {
        const int p = tile * 16 + row;
        const bool valid = p < P;
        Pix q{};
        if (valid) q = locate(im, p);
        f32x4 acc0 = {0.0f, 0.0f, 0.0f, 0.0f}, acc1 = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kc = 0; kc < KC; ++kc) {
            const int k0 = kc * 32 + kq * 8;
            const int t = k0 / CIN, ci = k0 % CIN;
            bf16x8 a;
            uint4 raw = make_uint4(0u, 0u, 0u, 0u);
            if (valid && t < 9) {
                const int s = tap_pixel(q, t, dil);
                if (s >= 0) raw = *reinterpret_cast<const uint4*>(in + (size_t)s * CIN + ci);
            }
            a = __builtin_bit_cast(bf16x8, raw);
            acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw[kc][0], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw[kc][1], acc1, 0, 0, 0);
        }
        // D[i][j]: lane = j + 16 (i / 4), element i % 4; i = pixel row, j = channel
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            const f32x4 acc = nt == 0 ? acc0 : acc1;
            const int co = nt * 16 + row;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int pp = tile * 16 + kq * 4 + e;
                if (pp >= P) continue;
                float v = acc[e];
                if (MODE == 0) {
                    v = fmaxf(v + bias[co], 0.0f);
                } else {
                    v = from_bf16(mask[(size_t)pp * kNnHidden + co]) > 0.0f ? v : 0.0f;
                }
                out[(size_t)pp * kNnHidden + co] = to_bf16(v);
            }
        }
    }
}

// 1x1 head (FP32) + output; with a target also the loss, the head's
// gradient partial sums (per block: 96 W + 3 b + loss) and the ReLU-masked
// gradient of h3.
__global__ __launch_bounds__(256) void k_head(int P, const uint16_t* h3, const float* head,
    const float4* beauty, const float4* target, float4* out, uint16_t* dh3, float* partial) {
    __shared__ float red[256];
    const int p = blockIdx.x * 256 + threadIdx.x;
    const bool valid = p < P;
    float h[kNnHidden];
    float dout[3] = {0.0f, 0.0f, 0.0f};
    float loss = 0.0f;
    if (valid) {
        const uint4* src = reinterpret_cast<const uint4*>(h3 + (size_t)p * kNnHidden);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const uint4 u = src[v];
            const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                h[v * 8 + 2 * k] = from_bf16((uint16_t)(w[k] & 0xFFFFu));
                h[v * 8 + 2 * k + 1] = from_bf16((uint16_t)(w[k] >> 16));
            }
        }
        const float4 b4 = beauty[p];
        const float base[3] = {log1pf(fmaxf(b4.x, 0.0f)), log1pf(fmaxf(b4.y, 0.0f)), log1pf(fmaxf(b4.z, 0.0f))};
        float o[3];
        const float scale = 1.0f / (3.0f * (float)P);
        float tl[3] = {0.0f, 0.0f, 0.0f};
        if (target) {
            const float4 t4 = target[p];
            tl[0] = log1pf(fmaxf(t4.x, 0.0f));
            tl[1] = log1pf(fmaxf(t4.y, 0.0f));
            tl[2] = log1pf(fmaxf(t4.z, 0.0f));
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            float y = head[3 * kNnHidden + c];
#pragma unroll
            for (int k = 0; k < kNnHidden; ++k) y = fmaf(head[c * kNnHidden + k], h[k], y);
            const float pred = base[c] + y;
            o[c] = fmaxf(expm1f(pred), 0.0f);
            if (target) {
                const float e = pred - tl[c];
                loss += e * e * scale;
                dout[c] = 2.0f * e * scale;
            }
        }
        if (out) out[p] = make_float4(o[0], o[1], o[2], 1.0f);
        if (target) {
            uint32_t packed[kNnHidden / 2];
#pragma unroll
            for (int k = 0; k < kNnHidden; k += 2) {
                float d0 = head[k] * dout[0] + head[kNnHidden + k] * dout[1] + head[2 * kNnHidden + k] * dout[2];
                float d1 = head[k + 1] * dout[0] + head[kNnHidden + k + 1] * dout[1]
                           + head[2 * kNnHidden + k + 1] * dout[2];
                d0 = h[k] > 0.0f ? d0 : 0.0f;
                d1 = h[k + 1] > 0.0f ? d1 : 0.0f;
                packed[k / 2] = (uint32_t)to_bf16(d0) | ((uint32_t)to_bf16(d1) << 16);
            }
            uint4* dst = reinterpret_cast<uint4*>(dh3 + (size_t)p * kNnHidden);
#pragma unroll
            for (int v = 0; v < 4; ++v)
                dst[v] = make_uint4(packed[4 * v], packed[4 * v + 1], packed[4 * v + 2], packed[4 * v + 3]);
        }
    }
    if (!target) return;
    // block sums of the head gradient: 96 W, 3 b, loss
    for (int i = 0; i < kNnHeadParams + 1; ++i) {
        float v = 0.0f;
        if (valid) {
            if (i < 3 * kNnHidden) v = dout[i / kNnHidden] * h[i % kNnHidden];
            else if (i < kNnHeadParams) v = dout[i - 3 * kNnHidden];
            else v = loss;
        }
        red[threadIdx.x] = v;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) partial[(size_t)blockIdx.x * (kNnHeadParams + 1) + i] = red[0];
        __syncthreads();
    }
}

// Weight gradient of one conv layer as a GEMM with K = pixels:
//   G[co][n] = sum_p dpre[p][co] * X[p + off(n / CIN)][n % CIN]   (n < 9 CIN)
//   G[co][9 CIN] = sum_p dpre[p][co]                              (bias)
// Block b covers pixels [b * kNnGradPixels, ...); wave w owns output tiles
// w, w + 4, ...; writes G to partial[b].
template <int CIN>
__global__ __launch_bounds__(256) void k_wgrad(NnImages im, int P, int dil, const uint16_t* dpre,
    const uint16_t* x, float* partial) {
    constexpr int NT = (9 * CIN) / 16 + 1;  // + bias tile
    constexpr int TILES = 2 * NT;
    constexpr int MAXT = (TILES + 3) / 4;
    constexpr int NW = NT * 16;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int col = lane & 15;
    const int kq = lane >> 4;
    f32x4 acc[MAXT];
#pragma unroll
    for (int i = 0; i < MAXT; ++i) acc[i] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const int p0 = blockIdx.x * kNnGradPixels;
    const int p1 = min(P, p0 + kNnGradPixels);
    for (int c0 = p0; c0 < p1; c0 += 32) {
        // this lane's 8 pixels of the 32-pixel chunk
        int pix[8];
        Pix q[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            pix[j] = c0 + kq * 8 + j;
            if (pix[j] < p1) q[j] = locate(im, pix[j]);
        }
#pragma unroll
        for (int i = 0; i < MAXT; ++i) {
            const int tile = wave + 4 * i;
            if (tile >= TILES) break;
            const int mt = tile / NT, nt = tile % NT;
            uint16_t av[8], bv[8];
            const int co = mt * 16 + col;
            const int n = nt * 16 + col;
            const int t = n / CIN, ci = n % CIN;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                av[j] = 0;
                bv[j] = 0;
                if (pix[j] >= p1) continue;
                av[j] = dpre[(size_t)pix[j] * kNnHidden + co];
                if (t < 9) {
                    const int s = tap_pixel(q[j], t, dil);
                    if (s >= 0) bv[j] = x[(size_t)s * CIN + ci];
                } else if (ci == 0) {
                    bv[j] = 0x3F80;  // 1.0: the bias column
                }
            }
            bf16x8 a, b;
            uint4 ra, rb;
            ra.x = (uint32_t)av[0] | ((uint32_t)av[1] << 16);
            ra.y = (uint32_t)av[2] | ((uint32_t)av[3] << 16);
            ra.z = (uint32_t)av[4] | ((uint32_t)av[5] << 16);
            ra.w = (uint32_t)av[6] | ((uint32_t)av[7] << 16);
            rb.x = (uint32_t)bv[0] | ((uint32_t)bv[1] << 16);
            rb.y = (uint32_t)bv[2] | ((uint32_t)bv[3] << 16);
            rb.z = (uint32_t)bv[4] | ((uint32_t)bv[5] << 16);
            rb.w = (uint32_t)bv[6] | ((uint32_t)bv[7] << 16);
            a = __builtin_bit_cast(bf16x8, ra);
            b = __builtin_bit_cast(bf16x8, rb);
            acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[i], 0, 0, 0);
        }
    }
    float* dst = partial + (size_t)blockIdx.x * 32 * NW;
#pragma unroll
    for (int i = 0; i < MAXT; ++i) {
        const int tile = wave + 4 * i;
        if (tile >= TILES) break;
        const int mt = tile / NT, nt = tile % NT;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int co = mt * 16 + kq * 4 + e;  // D row
            const int n = nt * 16 + col;          // D column
            dst[(size_t)co * NW + n] = acc[i][e];
        }
    }
}

// sum the per-block partials of one conv layer into the gradient vector
// (W[co][t][ci] then b[co]) in block order
template <int CIN>
__global__ void k_wgrad_reduce(int blocks, const float* partial, float* grad) {
    constexpr int NW = ((9 * CIN) / 16 + 1) * 16;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= 32 * (9 * CIN + 1)) return;
    int co, n;
    if (i < 32 * 9 * CIN) {
        co = i / (9 * CIN);
        n = i % (9 * CIN);
    } else {
        co = i - 32 * 9 * CIN;
        n = 9 * CIN;
    }
    float s = 0.0f;
    for (int b = 0; b < blocks; ++b) s += partial[(size_t)b * 32 * NW + (size_t)co * NW + n];
    grad[i] = s;
}

__global__ void k_head_reduce(int blocks, const float* partial, float* grad, float* loss) {
    const int i = threadIdx.x;
    if (i > kNnHeadParams) return;
    float s = 0.0f;
    for (int b = 0; b < blocks; ++b) s += partial[(size_t)b * (kNnHeadParams + 1) + i];
    if (i < kNnHeadParams) grad[i] = s;
    else *loss = s;
}

__global__ void k_adam(int n, float* params, const float* grad, float* m, float* v, float lr,
    float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float g = grad[i];
    const float mi = b1 * m[i] + (1.0f - b1) * g;
    const float vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    params[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// pack FP32 parameters into the kernels' layouts: forward bf16 [32][KP]
// (k = tap * CIN + ci), input-gradient bf16 [CIN][KP'] (k' = tap' * 32 + co,
// value W[co][8 - tap'][ci]), FP32 bias and head
__global__ void k_pack(const float* params, NnPacked pk) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    for (int l = 0; l < kNnLayers; ++l) {
        const int cin = l == 0 ? kNnFeat : kNnHidden;
        const int kp = ((9 * cin + 31) / 32) * 32;
        const float* W = params + nn_layer_offset(l);
        if (i < 32 * kp) {
            const int co = i / kp, k = i % kp;
            const float v = k < 9 * cin ? W[(size_t)co * 9 * cin + k] : 0.0f;
            pk.fwd[l][i] = to_bf16(v);
        }
        if (l > 0 && i < cin * 9 * 32) {
            const int ci = i / (9 * 32), k = i % (9 * 32);
            const int tp = k / 32, co = k % 32;
            pk.bwd[l][i] = to_bf16(W[((size_t)co * 9 + (8 - tp)) * cin + ci]);
        }
        if (i < 32) pk.bias[l][i] = W[(size_t)32 * 9 * cin + i];
    }
    if (i < kNnHeadParams) pk.head[i] = params[nn_head_offset() + i];
}

inline int cdiv(int a, int b) { return (a + b - 1) / b; }

}  // namespace

hipError_t nn_pack(const float* params, const NnPacked& pk, hipStream_t st) {
    k_pack<<<cdiv(32 * 288, 256), 256, 0, st>>>(params, pk);
    return hipGetLastError();
}

hipError_t nn_forward(const NnImages& im, int P, const NnPacked& pk, const NnScratch& w,
    const float4* beauty, const float4* normal, const float4* albedo, const float* depth,
    const float4* target, float4* out, int conv_grid, hipStream_t st) {
    if (P <= 0) return hipSuccess;
    k_features<<<cdiv(P, 256), 256, 0, st>>>(P, beauty, normal, albedo, depth, w.feat);
    const int dil[kNnLayers] = {1, 2, 4, 8};
    k_conv<kNnFeat, 0><<<conv_grid, 256, 0, st>>>(im, P, dil[0], w.feat, pk.fwd[0], pk.bias[0],
        nullptr, w.act[0]);
    for (int l = 1; l < kNnLayers; ++l) {
        k_conv<kNnHidden, 0><<<conv_grid, 256, 0, st>>>(im, P, dil[l], w.act[l - 1], pk.fwd[l],
            pk.bias[l], nullptr, w.act[l]);
    }
    k_head<<<cdiv(P, 256), 256, 0, st>>>(P, w.act[kNnLayers - 1], pk.head, beauty, target, out,
        w.dact[0], w.head_partial);
    return hipGetLastError();
}

hipError_t nn_backward(const NnImages& im, int P, const NnPacked& pk, const NnScratch& w,
    float* grad, float* loss, int conv_grid, hipStream_t st) {
    if (P <= 0) return hipSuccess;
    const int dil[kNnLayers] = {1, 2, 4, 8};
    const int hblocks = cdiv(P, 256);
    k_head_reduce<<<1, 128, 0, st>>>(hblocks, w.head_partial, grad + nn_head_offset(), loss);
    const int gblocks = cdiv(P, kNnGradPixels);
    int cur = 0;  // dact[cur] holds the pre-activation gradient of layer l
    for (int l = kNnLayers - 1; l >= 0; --l) {
        const uint16_t* xin = l == 0 ? w.feat : w.act[l - 1];
        if (l == 0) {
            k_wgrad<kNnFeat><<<gblocks, 256, 0, st>>>(im, P, dil[l], w.dact[cur], xin, w.grad_partial);
            k_wgrad_reduce<kNnFeat><<<cdiv(32 * (9 * kNnFeat + 1), 256), 256, 0, st>>>(gblocks,
                w.grad_partial, grad + nn_layer_offset(l));
        } else {
            k_wgrad<kNnHidden><<<gblocks, 256, 0, st>>>(im, P, dil[l], w.dact[cur], xin, w.grad_partial);
            k_wgrad_reduce<kNnHidden><<<cdiv(32 * (9 * kNnHidden + 1), 256), 256, 0, st>>>(gblocks,
                w.grad_partial, grad + nn_layer_offset(l));
            k_conv<kNnHidden, 1><<<conv_grid, 256, 0, st>>>(im, P, dil[l], w.dact[cur], pk.bwd[l],
                nullptr, w.act[l - 1], w.dact[cur ^ 1]);
            cur ^= 1;
        }
    }
    return hipGetLastError();
}

hipError_t nn_adam(int n, float* params, const float* grad, float* m, float* v, float lr, int step,
    hipStream_t st) {
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    k_adam<<<cdiv(n, 256), 256, 0, st>>>(n, params, grad, m, v, lr, b1, b2, 1e-8f, c1, c2);
    return hipGetLastError();
}

}  // namespace rtk
