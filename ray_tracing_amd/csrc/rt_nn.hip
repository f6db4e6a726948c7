// rt_nn.hip -- the learned denoiser (4-level encoder-decoder, see rt_nn.h):
// BF16 MFMA inference and training kernels with FP32 accumulation (gfx950
// v_mfma_f32_16x16x32_bf16).
//
// Every 3x3 convolution is one implicit GEMM: M = 16 pixels per MFMA tile,
// N = output channels (16-wide tiles), K = 9 taps x input channels in chunks
// of 32 (channel counts are multiples of 8, so an 8-wide K group never
// straddles a tap or the boundary between the two inputs of a decoder
// convolution).  A decoder convolution reads its second input one level
// coarser at the parent pixel (nearest upsampling folded into the load).
// The input gradient of a convolution is the same kernel on the output
// gradient with transposed, tap-flipped weights; it is written raw (bf16),
// and small kernels apply the ReLU masks, add the skip and pooled paths and
// sum the upsampled path back to the coarser level.  Weight gradients are
// GEMMs with K = pixels, written as per-block partial sums reduced in a fixed
// order (deterministic).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "rt_bf16.h"
#include "rt_nn.h"

namespace rtk {
namespace {

struct Pix {
    int img, base, w, h, x, y;
};

__device__ __forceinline__ Pix locate(const NnImages& im, int p) {
    int i = 0;
    while (i + 1 < im.count && p >= im.offset[i + 1]) ++i;
    Pix r;
    r.img = i;
    r.base = im.offset[i];
    r.w = im.width[i];
    r.h = (im.offset[i + 1] - im.offset[i]) / r.w;
    const int local = p - r.base;
    r.y = local / r.w;
    r.x = local - r.y * r.w;
    return r;
}

__device__ __forceinline__ int image_height(const NnImages& im, int i) {
    return (im.offset[i + 1] - im.offset[i]) / im.width[i];
}

// index of the parent (one level coarser) of pixel (x, y) of image i
__device__ __forceinline__ int parent_index(const NnImages& pim, int i, int x, int y) {
    return pim.offset[i] + (y >> 1) * pim.width[i] + (x >> 1);
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

struct ConvArgs {
    NnImages im;          // this level
    NnImages pim;         // one level coarser (second input)
    int P;                // pixels at this level
    int Pp;               // pixels one level coarser
    const uint16_t* A;    // [P][C0]
    const uint16_t* B;    // [P coarser][C1]
    const uint16_t* W;    // bf16 [COUT][kp]
    int kp;
    const float* bias;
    uint16_t* out;        // [P][COUT]
};

// Raw buffer resource over `bytes` bytes (gfx9 descriptor word 3: 32-bit
// data format); a load at an offset past the end returns zeros, which is how
// taps outside the image and padding channels read as zero without a branch.
constexpr uint32_t kBufWord3 = 0x00020000u;
constexpr int kBufOob = 0x7FFFF000;  // beyond every buffer, also plus an immediate offset < 4096 (sizes are checked < 2^31 - 2^12 bytes)

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, (int)kBufWord3);
}

__device__ __forceinline__ bf16x8 buf_load16(__amdgpu_buffer_rsrc_t r, int off) {
    return __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}

// pixel tiles (of 16) per wave: each weight fragment feeds MT MFMAs
template <int COUT>
constexpr int conv_mt() { return COUT / 16 <= 4 ? 4 : 2; }

// One wave computes out[16 MT pixels][COUT] as the transposed GEMM
// out^T = W X^T: the MFMA A operand is the weights (row = output channel,
// 8 consecutive K per lane, one 16-byte load), the B operand the input
// (column = pixel, 8 channels of one tap per lane).  K runs tap by tap, each
// tap's channels padded to a multiple of 32, so every MFMA step has one
// compile-time tap: a lane's input address is its pixel's precomputed base
// plus a constant, and the tap's in-image bit selects it or an out-of-range
// offset (zeros).  Decoder convolutions read channels >= C0 at the parent
// pixel of the coarser level (nearest upsampling folded into the load).
// A lane ends with 4 consecutive output channels of one pixel and stores
// them as 8 bytes.  MODE 0: out = relu(acc + bias); MODE 1: out = acc.
template <int C0, int C1, int COUT, int MODE>
__global__ __launch_bounds__(256) void k_uconv(ConvArgs a) {
    constexpr int CIN = C0 + C1;
    constexpr int TS = nn_tap_stride(CIN);
    constexpr int KT = TS / 32;
    constexpr int KP = 9 * TS;
    constexpr int NT = COUT / 16;
    constexpr int MT = conv_mt<COUT>();
    static_assert(CIN % 8 == 0 && C0 % 8 == 0 && COUT % 16 == 0, "channel layout");
    const int lane = threadIdx.x & 63;
    const int r16 = lane & 15;
    const int kq = lane >> 4;
    const int waves = gridDim.x * 4;
    const int groups = (a.P + 16 * MT - 1) / (16 * MT);
    const __amdgpu_buffer_rsrc_t ra = buf_rsrc(a.A, (uint32_t)a.P * C0 * 2u);
    const __amdgpu_buffer_rsrc_t rb = buf_rsrc(C1 > 0 ? a.B : a.A, C1 > 0 ? (uint32_t)a.Pp * C1 * 2u : 0u);
    const uint16_t* wrow = a.W + (size_t)r16 * KP + kq * 8;
    for (int g = blockIdx.x * 4 + (threadIdx.x >> 6); g < groups; g += waves) {
        // per pixel tile: byte offset of this lane's pixel and channel octet,
        // row pitch, in-image bits of the 9 taps; decoder: the coarser image's
        // base offset and row pitch and the pixel's coordinates
        int cA[MT], rowA[MT], cB[MT], px_[MT], py_[MT], pwC[MT];
        uint32_t mask[MT];
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            const int p = (g * MT + m) * 16 + r16;
            mask[m] = 0u;
            cA[m] = rowA[m] = cB[m] = px_[m] = py_[m] = pwC[m] = 0;
            if (p < a.P) {
                const Pix q = locate(a.im, p);
                cA[m] = ((q.base + q.y * q.w + q.x) * C0 + kq * 8) * 2;
                rowA[m] = q.w * C0 * 2;
#pragma unroll
                for (int t = 0; t < 9; ++t) {
                    const int sx = q.x + t % 3 - 1, sy = q.y + t / 3 - 1;
                    mask[m] |= (sx >= 0 && sx < q.w && sy >= 0 && sy < q.h) ? (1u << t) : 0u;
                }
                if (C1 > 0) {
                    cB[m] = (a.pim.offset[q.img] * C1 + kq * 8) * 2;
                    pwC[m] = a.pim.width[q.img] * C1 * 2;
                    px_[m] = q.x;
                    py_[m] = q.y;
                }
            }
        }
        f32x4 acc[MT][NT];
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) acc[m][nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        // per-lane channel validity of each K step (padding past CIN reads
        // zero) and, for a step straddling the two inputs, which one it reads
        bool ch_ok[KT], in_a[KT];
#pragma unroll
        for (int c = 0; c < KT; ++c) {
            ch_ok[c] = c * 32 + kq * 8 < CIN;
            in_a[c] = c * 32 + kq * 8 < C0;
        }
#pragma unroll 1
        for (int t = 0; t < 9; ++t) {
            const int dy = t / 3 - 1, dx = t % 3 - 1;
            // this tap's input base offsets (out-of-range when outside the image)
            int baseA[MT], baseB[MT];
#pragma unroll
            for (int m = 0; m < MT; ++m) {
                const bool tap_ok = ((mask[m] >> t) & 1u) != 0u;
                const int rowoff = dy < 0 ? -rowA[m] : dy > 0 ? rowA[m] : 0;
                baseA[m] = tap_ok ? cA[m] + rowoff + dx * (C0 * 2) : kBufOob;
                baseB[m] = kBufOob;
                if (C1 > 0) {
                    // parent of (x + dx, y + dy) at the coarser level
                    const int prow = ((py_[m] + dy) >> 1) * pwC[m];
                    const int pcol = ((px_[m] + dx) >> 1) * (C1 * 2);
                    baseB[m] = tap_ok ? cB[m] + prow + pcol : kBufOob;
                }
            }
            const uint16_t* wt = wrow + t * TS;
#pragma unroll
            for (int c = 0; c < KT; ++c) {
                const int cb = c * 32;   // channel of lane group kq = 0
                bf16x8 wv[NT];
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                    wv[nt] = *reinterpret_cast<const bf16x8*>(wt + (size_t)nt * 16 * KP + cb);
                const bool full = cb + 32 <= CIN;
                bf16x8 xv[MT];
#pragma unroll
                for (int m = 0; m < MT; ++m) {
                    if (C1 == 0 || cb + 32 <= C0) {
                        xv[m] = buf_load16(ra, (full || ch_ok[c] ? baseA[m] : kBufOob) + cb * 2);
                    } else if (cb >= C0) {
                        xv[m] = buf_load16(rb, (full || ch_ok[c] ? baseB[m] : kBufOob) + (cb - C0) * 2);
                    } else {
                        const bf16x8 xa = buf_load16(ra, (in_a[c] ? baseA[m] : kBufOob) + cb * 2);
                        const bf16x8 xb = buf_load16(rb, (!in_a[c] && ch_ok[c] ? baseB[m] : kBufOob) + (cb - C0) * 2);
                        xv[m] = __builtin_bit_cast(bf16x8, __builtin_bit_cast(uint4, xa) | __builtin_bit_cast(uint4, xb));
                    }
                }
#pragma unroll
                for (int m = 0; m < MT; ++m)
#pragma unroll
                    for (int nt = 0; nt < NT; ++nt)
                        acc[m][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wv[nt], xv[m], acc[m][nt], 0, 0, 0);
            }
        }
        // D[i][j]: i = output channel 4 kq + e of the tile, j = pixel r16
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            const int p = (g * MT + m) * 16 + r16;
            if (p >= a.P) continue;
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) {
                const int co = nt * 16 + kq * 4;
                float v[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    v[e] = acc[m][nt][e];
                    if (MODE == 0) v[e] = fmaxf(v[e] + a.bias[co + e], 0.0f);
                }
                const uint2 packed = make_uint2((uint32_t)to_bf16(v[0]) | ((uint32_t)to_bf16(v[1]) << 16),
                                                (uint32_t)to_bf16(v[2]) | ((uint32_t)to_bf16(v[3]) << 16));
                *reinterpret_cast<uint2*>(a.out + (size_t)p * COUT + co) = packed;
            }
        }
    }
}

// mean of the pixels of level l that a level-(l+1) pixel covers; one thread
// per (coarse pixel, 8 channels): 16-byte loads and stores
__global__ __launch_bounds__(256) void k_pool(NnImages im, NnImages pim, int Pp, int C,
    const uint16_t* in, uint16_t* out) {
    const int c8 = C / 8;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= Pp * c8) return;
    const int pp = i / c8, c = (i - pp * c8) * 8;
    const Pix q = locate(pim, pp);
    const int w = im.width[q.img], h = image_height(im, q.img), base = im.offset[q.img];
    float s[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    int n = 0;
    for (int dy = 0; dy < 2; ++dy) {
        for (int dx = 0; dx < 2; ++dx) {
            const int x = 2 * q.x + dx, y = 2 * q.y + dy;
            if (x < w && y < h) {
                const uint4 raw = *reinterpret_cast<const uint4*>(in + (size_t)(base + y * w + x) * C + c);
                const uint32_t v[4] = {raw.x, raw.y, raw.z, raw.w};
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    s[2 * j] += from_bf16((uint16_t)(v[j] & 0xFFFFu));
                    s[2 * j + 1] += from_bf16((uint16_t)(v[j] >> 16));
                }
                ++n;
            }
        }
    }
    const float inv = (float)n;
    uint32_t o[4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
        o[j] = (uint32_t)to_bf16(s[2 * j] / inv) | ((uint32_t)to_bf16(s[2 * j + 1] / inv) << 16);
    *reinterpret_cast<uint4*>(out + (size_t)pp * C + c) = make_uint4(o[0], o[1], o[2], o[3]);
}

// dz = r where act > 0
__global__ __launch_bounds__(256) void k_mask(size_t n, const uint16_t* r, const uint16_t* act,
    uint16_t* dz) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    dz[i] = from_bf16(act[i]) > 0.0f ? r[i] : (uint16_t)0;
}

// dz at level l+1 = sum over the covered level-l pixels of r[:, off .. off + C)
// (the upsampled input's gradient), where act > 0
__global__ __launch_bounds__(256) void k_up_combine(NnImages im, NnImages pim, int Pp, int C,
    const uint16_t* r, int cr, int off, const uint16_t* act, uint16_t* dz) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= Pp * C) return;
    const int pp = i / C, c = i - pp * C;
    const Pix q = locate(pim, pp);
    const int w = im.width[q.img], h = image_height(im, q.img), base = im.offset[q.img];
    float s = 0.0f;
    for (int dy = 0; dy < 2; ++dy) {
        for (int dx = 0; dx < 2; ++dx) {
            const int x = 2 * q.x + dx, y = 2 * q.y + dy;
            if (x < w && y < h) s += from_bf16(r[(size_t)(base + y * w + x) * cr + off + c]);
        }
    }
    dz[(size_t)pp * C + c] = from_bf16(act[(size_t)pp * C + c]) > 0.0f ? to_bf16(s) : (uint16_t)0;
}

// dz at level l = (skip gradient rs[:, 0..C) + pooled-path gradient rp at the
// parent / pixels the parent covers), where act > 0
__global__ __launch_bounds__(256) void k_skip_combine(NnImages im, NnImages pim, int P, int C,
    const uint16_t* rs, int cs, const uint16_t* rp, const uint16_t* act, uint16_t* dz) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= P * C) return;
    const int p = i / C, c = i - p * C;
    const Pix q = locate(im, p);
    const int px = q.x >> 1, py = q.y >> 1;
    const int n = min(2, q.w - 2 * px) * min(2, q.h - 2 * py);
    const int pp = pim.offset[q.img] + py * pim.width[q.img] + px;
    const float v = from_bf16(rs[(size_t)p * cs + c]) + from_bf16(rp[(size_t)pp * C + c]) / (float)n;
    dz[(size_t)p * C + c] = from_bf16(act[(size_t)p * C + c]) > 0.0f ? to_bf16(v) : (uint16_t)0;
}

// Per-channel weight of the log-space squared error: with e = log1p(y) -
// log1p(t) small, y - t ~ (1 + t) e, so w e^2 with w = (1 + t)^2 / (t^2 + 1e-2)
// is the relative squared error (y - t)^2 / (t^2 + 1e-2) to first order while
// the gradient of a large error (a firefly) stays linear in e.
__device__ __forceinline__ float loss_weight(float t) {
    const float a = 1.0f + t;
    return a * a / (t * t + 1e-2f);
}

// 1x1 head (FP32) + output; with a target also the loss, the head's
// gradient partial sums (per block: 96 W + 3 b + loss) and the ReLU-masked
// gradient of d0.
__global__ __launch_bounds__(256) void k_head(int P, const uint16_t* h3, const float* head,
    const float4* beauty, const float4* target, float4* out, uint16_t* dh3, float* partial) {
    __shared__ float red[256];
    const int p = blockIdx.x * 256 + threadIdx.x;
    const bool valid = p < P;
    float h[kNnOut];
    float dout[3] = {0.0f, 0.0f, 0.0f};
    float loss = 0.0f;
    if (valid) {
        const uint4* src = reinterpret_cast<const uint4*>(h3 + (size_t)p * kNnOut);
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
        float wt[3] = {0.0f, 0.0f, 0.0f};
        if (target) {
            const float4 t4 = target[p];
            const float t[3] = {fmaxf(t4.x, 0.0f), fmaxf(t4.y, 0.0f), fmaxf(t4.z, 0.0f)};
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                tl[c] = log1pf(t[c]);
                wt[c] = loss_weight(t[c]);
            }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            float y = head[3 * kNnOut + c];
#pragma unroll
            for (int k = 0; k < kNnOut; ++k) y = fmaf(head[c * kNnOut + k], h[k], y);
            const float pred = base[c] + y;
            o[c] = fmaxf(expm1f(pred), 0.0f);
            if (target) {
                const float e = pred - tl[c];
                loss += wt[c] * e * e * scale;
                dout[c] = 2.0f * wt[c] * e * scale;
            }
        }
        if (out) out[p] = make_float4(o[0], o[1], o[2], 1.0f);
        if (target) {
            uint32_t packed[kNnOut / 2];
#pragma unroll
            for (int k = 0; k < kNnOut; k += 2) {
                float d0 = head[k] * dout[0] + head[kNnOut + k] * dout[1] + head[2 * kNnOut + k] * dout[2];
                float d1 = head[k + 1] * dout[0] + head[kNnOut + k + 1] * dout[1]
                           + head[2 * kNnOut + k + 1] * dout[2];
                d0 = h[k] > 0.0f ? d0 : 0.0f;
                d1 = h[k + 1] > 0.0f ? d1 : 0.0f;
                packed[k / 2] = (uint32_t)to_bf16(d0) | ((uint32_t)to_bf16(d1) << 16);
            }
            uint4* dst = reinterpret_cast<uint4*>(dh3 + (size_t)p * kNnOut);
#pragma unroll
            for (int v = 0; v < 4; ++v)
                dst[v] = make_uint4(packed[4 * v], packed[4 * v + 1], packed[4 * v + 2], packed[4 * v + 3]);
        }
    }
    if (!target) return;
    for (int i = 0; i < kNnHeadParams + 1; ++i) {
        float v = 0.0f;
        if (valid) {
            if (i < 3 * kNnOut) v = dout[i / kNnOut] * h[i % kNnOut];
            else if (i < kNnHeadParams) v = dout[i - 3 * kNnOut];
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

struct WgradArgs {
    NnImages im, pim;
    int P;
    const uint16_t* dz;   // [P][cout]
    const uint16_t* A;    // [P][C0]
    const uint16_t* B;    // [P coarser][C1]
    float* partial;       // [chunks][cout][NW * 16]
};

constexpr int kWgStep = 64;            // pixels per LDS stage (two MFMA K slices)
constexpr int kWgRow = kWgStep + 8;    // LDS row stride in bf16: 144 B, 16-byte aligned rows

// scatter 8 channels of one pixel into an LDS [channel][pixel] tile
__device__ __forceinline__ void lds_scatter8(uint16_t* s, int c, int px, uint4 raw) {
    const uint32_t w[4] = {raw.x, raw.y, raw.z, raw.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        s[(c + 2 * j) * kWgRow + px] = (uint16_t)(w[j] & 0xFFFFu);
        s[(c + 2 * j + 1) * kWgRow + px] = (uint16_t)(w[j] >> 16);
    }
}

// Weight gradient of one convolution as GEMMs with K = pixels:
//   G[co][t CIN + ci] = sum_p dz[p][co] X[p + off(t)][ci]   (block row t < 9)
//   G[co][9 CIN]      = sum_p dz[p][co]                      (block row t = 9, the bias)
// Block chunk * 10 + t covers pixels [chunk * kNnGradPixels, ...).  Per stage of
// 64 pixels it copies dz and the tap-shifted input into LDS transposed
// ([channel][pixel]), so that a lane reads the 8 consecutive pixels of one
// channel an MFMA operand needs with one 16-byte LDS load; wave w owns the
// (cout / 16) x (CIN / 16) output tiles w, w + 4, ...  Partial sums per chunk
// are reduced in chunk order afterwards (deterministic).
template <int C0, int C1, int COUT>
__global__ __launch_bounds__(256) void k_wgrad(WgradArgs a) {
    constexpr int CIN = C0 + C1;
    constexpr int MT = COUT / 16, NT = CIN / 16;
    constexpr int TILES = MT * NT;
    constexpr int TPW = (TILES + 3) / 4;
    constexpr int NWC = ((9 * CIN) / 16 + 1) * 16;
    static_assert(CIN % 16 == 0 && C0 % 8 == 0 && COUT % 16 == 0, "channel layout");
    __shared__ __align__(16) uint16_t sdz[COUT * kWgRow];
    __shared__ __align__(16) uint16_t sx[CIN * kWgRow];
    __shared__ int src_a[kWgStep], src_b[kWgStep];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int r16 = lane & 15, kq = lane >> 4;
    const int chunk = (int)blockIdx.x / 10, t = (int)blockIdx.x - chunk * 10;
    const bool bias = t == 9;
    const int dx = t % 3 - 1, dy = t / 3 - 1;
    const int tiles = bias ? MT : TILES;
    f32x4 acc[TPW];
#pragma unroll
    for (int i = 0; i < TPW; ++i) acc[i] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const int p0 = chunk * kNnGradPixels;
    const int p1 = min(a.P, p0 + kNnGradPixels);
    const bf16x8 ones = __builtin_bit_cast(bf16x8, make_uint4(0x3F803F80u, 0x3F803F80u, 0x3F803F80u, 0x3F803F80u));
    for (int s0 = p0; s0 < p1; s0 += kWgStep) {
        if (!bias && tid < kWgStep) {
            const int p = s0 + tid;
            int ia = -1, ib = -1;
            if (p < p1) {
                const Pix q = locate(a.im, p);
                const int sx_ = q.x + dx, sy_ = q.y + dy;
                if (sx_ >= 0 && sx_ < q.w && sy_ >= 0 && sy_ < q.h) {
                    ia = q.base + sy_ * q.w + sx_;
                    if (C1 > 0) ib = parent_index(a.pim, q.img, sx_, sy_);
                }
            }
            src_a[tid] = ia;
            src_b[tid] = ib;
        }
        for (int u = tid; u < kWgStep * (COUT / 8); u += 256) {
            const int px = u / (COUT / 8), c = (u - px * (COUT / 8)) * 8;
            const int p = s0 + px;
            uint4 raw = make_uint4(0u, 0u, 0u, 0u);
            if (p < p1) raw = *reinterpret_cast<const uint4*>(a.dz + (size_t)p * COUT + c);
            lds_scatter8(sdz, c, px, raw);
        }
        __syncthreads();
        if (!bias) {
            for (int u = tid; u < kWgStep * (CIN / 8); u += 256) {
                const int px = u / (CIN / 8), c = (u - px * (CIN / 8)) * 8;
                uint4 raw = make_uint4(0u, 0u, 0u, 0u);
                if (C1 == 0 || c < C0) {
                    const int i = src_a[px];
                    if (i >= 0) raw = *reinterpret_cast<const uint4*>(a.A + (size_t)i * C0 + c);
                } else {
                    const int i = src_b[px];
                    if (i >= 0) raw = *reinterpret_cast<const uint4*>(a.B + (size_t)i * C1 + (c - C0));
                }
                lds_scatter8(sx, c, px, raw);
            }
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < kWgStep / 32; ++ks) {
#pragma unroll
            for (int i = 0; i < TPW; ++i) {
                const int tile = wave + 4 * i;
                if (tile < tiles) {
                    const int mt = bias ? tile : tile / NT, nt = bias ? 0 : tile - (tile / NT) * NT;
                    const bf16x8 av = *reinterpret_cast<const bf16x8*>(sdz + (mt * 16 + r16) * kWgRow + ks * 32 + kq * 8);
                    const bf16x8 bv = bias ? ones
                                           : *reinterpret_cast<const bf16x8*>(sx + (nt * 16 + r16) * kWgRow + ks * 32 + kq * 8);
                    acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[i], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    float* dst = a.partial + (size_t)chunk * COUT * NWC;
#pragma unroll
    for (int i = 0; i < TPW; ++i) {
        const int tile = wave + 4 * i;
        if (tile >= tiles) continue;
        const int mt = bias ? tile : tile / NT, nt = bias ? 0 : tile - (tile / NT) * NT;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int co = mt * 16 + kq * 4 + e;   // D row
            if (bias) {
                if (r16 == 0) dst[(size_t)co * NWC + 9 * CIN] = acc[i][e];
            } else {
                dst[(size_t)co * NWC + t * CIN + nt * 16 + r16] = acc[i][e];
            }
        }
    }
}

// sum the chunks' partials of one convolution into the gradient vector
// (W[co][tap][ci] then b[co]) in chunk order
__global__ void k_uwgrad_reduce(int chunks, int cout, int cin, const float* partial, float* grad) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int nw16 = ((9 * cin) / 16 + 1) * 16;
    if (i >= cout * (9 * cin + 1)) return;
    int co, n;
    if (i < cout * 9 * cin) {
        co = i / (9 * cin);
        n = i - co * 9 * cin;
    } else {
        co = i - cout * 9 * cin;
        n = 9 * cin;
    }
    float s = 0.0f;
    for (int b = 0; b < chunks; ++b) s += partial[((size_t)b * cout + co) * nw16 + n];
    grad[i] = s;
}

// block i sums column i of the per-block head partials (fixed order)
__global__ __launch_bounds__(256) void k_head_reduce(int blocks, const float* partial, float* grad,
    float* loss) {
    __shared__ float red[256];
    const int i = blockIdx.x;
    float s = 0.0f;
    for (int b = threadIdx.x; b < blocks; b += 256) s += partial[(size_t)b * (kNnHeadParams + 1) + i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (i < kNnHeadParams) grad[i] = red[0];
        else *loss = red[0];
    }
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

// pack convolution l: forward bf16 [cout][kp] (k = tap * stride(cin) + ci),
// input-gradient bf16 [cin][kp'] (k' = tap' * stride(cout) + co, value
// W[co][8 - tap'][ci]), padding zero; FP32 bias
__global__ void k_pack_conv(const float* W, int cin, int cout, int kpf, int kpb, uint16_t* fwd,
    uint16_t* bwd, float* bias) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < cout * kpf) {
        const int ts = kpf / 9;
        const int co = i / kpf, k = i - co * kpf;
        const int t = k / ts, ci = k - t * ts;
        fwd[i] = ci < cin ? to_bf16(W[((size_t)co * 9 + t) * cin + ci]) : (uint16_t)0;
    }
    if (bwd && i < cin * kpb) {
        const int ts = kpb / 9;
        const int ci = i / kpb, k = i - ci * kpb;
        const int tp = k / ts, co = k - tp * ts;
        bwd[i] = co < cout ? to_bf16(W[((size_t)co * 9 + (8 - tp)) * cin + ci]) : (uint16_t)0;
    }
    if (i < cout) bias[i] = W[(size_t)cout * 9 * cin + i];
}

__global__ void k_copy(int n, const float* src, float* dst) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

inline int cdiv(int a, int b) { return (a + b - 1) / b; }

inline int level_pixels(const NnImages& im) { return im.offset[im.count]; }

// forward (MODE 0) or input-gradient (MODE 1) launch of convolution l
template <int C0, int C1, int COUT, int MODE>
void conv_launch(const ConvArgs& a, int grid, hipStream_t st) {
    const int groups = cdiv(a.P, 16 * conv_mt<COUT>());
    const int blocks = std::max(1, std::min(grid, cdiv(groups, 4)));
    k_uconv<C0, C1, COUT, MODE><<<blocks, 256, 0, st>>>(a);
}

// the convolutions address their inputs with 32-bit buffer offsets
bool conv_inputs_fit(const NnLevels& lv) {
    constexpr double kLimit = 2147479552.0;  // kBufOob
    for (int l = 0; l < kNnConvs; ++l) {
        const NnConv c = nn_conv(l);
        const double P = (double)lv.lv[c.level].offset[lv.lv[c.level].count];
        if (P * (double)std::max(nn_cin(l), c.cout) * 2.0 >= kLimit) return false;
    }
    return true;
}

void conv_forward(int l, const ConvArgs& a, int grid, hipStream_t st) {
    switch (l) {
        case 0: conv_launch<16, 0, 32, 0>(a, grid, st); break;
        case 1: conv_launch<32, 0, 32, 0>(a, grid, st); break;
        case 2: conv_launch<32, 0, 48, 0>(a, grid, st); break;
        case 3: conv_launch<48, 0, 48, 0>(a, grid, st); break;
        case 4: conv_launch<48, 0, 64, 0>(a, grid, st); break;
        case 5: conv_launch<64, 0, 64, 0>(a, grid, st); break;
        case 6: conv_launch<64, 0, 96, 0>(a, grid, st); break;
        case 7: conv_launch<96, 0, 96, 0>(a, grid, st); break;
        case 8: conv_launch<96, 0, 96, 0>(a, grid, st); break;
        case 9: conv_launch<64, 96, 64, 0>(a, grid, st); break;
        case 10: conv_launch<64, 0, 64, 0>(a, grid, st); break;
        case 11: conv_launch<48, 64, 48, 0>(a, grid, st); break;
        case 12: conv_launch<48, 0, 48, 0>(a, grid, st); break;
        case 13: conv_launch<32, 48, 32, 0>(a, grid, st); break;
        default: conv_launch<32, 0, 32, 0>(a, grid, st); break;
    }
}

// input gradient of convolution l: the output gradient (cout channels) through
// the transposed weights into cin raw channels
void conv_input_grad(int l, const ConvArgs& a, int grid, hipStream_t st) {
    switch (l) {
        case 1: conv_launch<32, 0, 32, 1>(a, grid, st); break;
        case 2: conv_launch<48, 0, 32, 1>(a, grid, st); break;
        case 3: conv_launch<48, 0, 48, 1>(a, grid, st); break;
        case 4: conv_launch<64, 0, 48, 1>(a, grid, st); break;
        case 5: conv_launch<64, 0, 64, 1>(a, grid, st); break;
        case 6: conv_launch<96, 0, 64, 1>(a, grid, st); break;
        case 7: conv_launch<96, 0, 96, 1>(a, grid, st); break;
        case 8: conv_launch<96, 0, 96, 1>(a, grid, st); break;
        case 9: conv_launch<64, 0, 160, 1>(a, grid, st); break;
        case 10: conv_launch<64, 0, 64, 1>(a, grid, st); break;
        case 11: conv_launch<48, 0, 112, 1>(a, grid, st); break;
        case 12: conv_launch<48, 0, 48, 1>(a, grid, st); break;
        case 13: conv_launch<32, 0, 80, 1>(a, grid, st); break;
        default: conv_launch<32, 0, 32, 1>(a, grid, st); break;
    }
}

template <int C0, int C1, int COUT>
void wgrad_launch(const WgradArgs& a, hipStream_t st) {
    k_wgrad<C0, C1, COUT><<<cdiv(a.P, kNnGradPixels) * 10, 256, 0, st>>>(a);
}

void conv_wgrad(int l, const WgradArgs& a, hipStream_t st) {
    switch (l) {
        case 0: wgrad_launch<16, 0, 32>(a, st); break;
        case 1: wgrad_launch<32, 0, 32>(a, st); break;
        case 2: wgrad_launch<32, 0, 48>(a, st); break;
        case 3: wgrad_launch<48, 0, 48>(a, st); break;
        case 4: wgrad_launch<48, 0, 64>(a, st); break;
        case 5: wgrad_launch<64, 0, 64>(a, st); break;
        case 6: wgrad_launch<64, 0, 96>(a, st); break;
        case 7: wgrad_launch<96, 0, 96>(a, st); break;
        case 8: wgrad_launch<96, 0, 96>(a, st); break;
        case 9: wgrad_launch<64, 96, 64>(a, st); break;
        case 10: wgrad_launch<64, 0, 64>(a, st); break;
        case 11: wgrad_launch<48, 64, 48>(a, st); break;
        case 12: wgrad_launch<48, 0, 48>(a, st); break;
        case 13: wgrad_launch<32, 48, 32>(a, st); break;
        default: wgrad_launch<32, 0, 32>(a, st); break;
    }
}

// the inputs of convolution l: A (this level) and B (one level coarser)
void conv_inputs(int l, const NnScratch& w, const uint16_t** A, const uint16_t** B) {
    *B = nullptr;
    switch (l) {
        case 0: *A = w.feat; break;
        case 2: *A = w.pool[1]; break;
        case 4: *A = w.pool[2]; break;
        case 6: *A = w.pool[3]; break;
        case 9: *A = w.act[5]; *B = w.act[8]; break;
        case 11: *A = w.act[3]; *B = w.act[10]; break;
        case 13: *A = w.act[1]; *B = w.act[12]; break;
        default: *A = w.act[l - 1]; break;
    }
}

ConvArgs conv_args(const NnLevels& lv, int l, const NnScratch& w) {
    const NnConv c = nn_conv(l);
    ConvArgs a{};
    a.im = lv.lv[c.level];
    a.pim = lv.lv[c.level < kNnLevels - 1 ? c.level + 1 : c.level];
    a.P = level_pixels(a.im);
    a.Pp = level_pixels(a.pim);
    conv_inputs(l, w, &a.A, &a.B);
    return a;
}

}  // namespace

NnLevels nn_levels(const NnImages& im) {
    NnLevels r{};
    r.lv[0] = im;
    for (int l = 1; l < kNnLevels; ++l) {
        const NnImages& u = r.lv[l - 1];
        NnImages& d = r.lv[l];
        d.count = u.count;
        int o = 0;
        for (int i = 0; i < u.count; ++i) {
            const int w = u.width[i], h = (u.offset[i + 1] - u.offset[i]) / u.width[i];
            d.offset[i] = o;
            d.width[i] = (w + 1) / 2;
            o += ((w + 1) / 2) * ((h + 1) / 2);
        }
        d.offset[u.count] = o;
    }
    return r;
}

size_t nn_grad_partial_floats(const NnLevels& lv) {
    size_t m = 0;
    for (int l = 0; l < kNnConvs; ++l) {
        const NnConv c = nn_conv(l);
        const int chunks = cdiv(level_pixels(lv.lv[c.level]), kNnGradPixels);
        const size_t nw16 = (size_t)((9 * nn_cin(l)) / 16 + 1) * 16;
        m = std::max(m, (size_t)chunks * (size_t)c.cout * nw16);
    }
    return m;
}

hipError_t nn_pack(const float* params, const NnPacked& pk, hipStream_t st) {
    for (int l = 0; l < kNnConvs; ++l) {
        const NnConv c = nn_conv(l);
        const int cin = nn_cin(l);
        const int n = std::max(c.cout * nn_kp_fwd(l), cin * nn_kp_bwd(l));
        k_pack_conv<<<cdiv(n, 256), 256, 0, st>>>(params + nn_layer_offset(l), cin, c.cout,
            nn_kp_fwd(l), nn_kp_bwd(l), pk.fwd[l], l > 0 ? pk.bwd[l] : nullptr, pk.bias[l]);
    }
    k_copy<<<1, 256, 0, st>>>(kNnHeadParams, params + nn_head_offset(), pk.head);
    return hipGetLastError();
}

hipError_t nn_forward(const NnLevels& lv, const NnPacked& pk, const NnScratch& w,
    const float4* beauty, const float4* normal, const float4* albedo, const float* depth,
    const float4* target, float4* out, int conv_grid, hipStream_t st) {
    const int P = level_pixels(lv.lv[0]);
    if (P <= 0) return hipSuccess;
    if (!conv_inputs_fit(lv)) return hipErrorInvalidValue;
    k_features<<<cdiv(P, 256), 256, 0, st>>>(P, beauty, normal, albedo, depth, w.feat);
    for (int l = 0; l < kNnConvs; ++l) {
        ConvArgs a = conv_args(lv, l, w);
        a.W = pk.fwd[l];
        a.kp = nn_kp_fwd(l);
        a.bias = pk.bias[l];
        a.out = w.act[l];
        conv_forward(l, a, conv_grid, st);
        if (l == 1 || l == 3 || l == 5) {
            // pooled e0 / e1 / e2 feed the next level
            const int lvl = nn_conv(l).level;
            const int C = nn_conv(l).cout;
            const int Pp = level_pixels(lv.lv[lvl + 1]);
            k_pool<<<cdiv(Pp * (C / 8), 256), 256, 0, st>>>(lv.lv[lvl], lv.lv[lvl + 1], Pp, C, w.act[l],
                w.pool[lvl + 1]);
        }
    }
    k_head<<<cdiv(P, 256), 256, 0, st>>>(P, w.act[kNnConvs - 1], pk.head, beauty, target, out,
        w.dz[kNnConvs - 1], w.head_partial);
    return hipGetLastError();
}

hipError_t nn_backward(const NnLevels& lv, const NnPacked& pk, const NnScratch& w, float* grad,
    float* loss, int conv_grid, hipStream_t st) {
    const int P = level_pixels(lv.lv[0]);
    if (P <= 0) return hipSuccess;
    k_head_reduce<<<kNnHeadParams + 1, 256, 0, st>>>(cdiv(P, 256), w.head_partial, grad + nn_head_offset(), loss);
    for (int l = kNnConvs - 1; l >= 0; --l) {
        const NnConv c = nn_conv(l);
        const int cin = nn_cin(l);
        const NnImages& im = lv.lv[c.level];
        const int Pl = level_pixels(im);
        // weight gradient
        WgradArgs g{};
        g.im = im;
        g.pim = lv.lv[c.level < kNnLevels - 1 ? c.level + 1 : c.level];
        g.P = Pl;
        g.dz = w.dz[l];
        conv_inputs(l, w, &g.A, &g.B);
        g.partial = w.grad_partial;
        conv_wgrad(l, g, st);
        k_uwgrad_reduce<<<cdiv(c.cout * (9 * cin + 1), 256), 256, 0, st>>>(
            cdiv(Pl, kNnGradPixels), c.cout, cin, w.grad_partial, grad + nn_layer_offset(l));
        if (l == 0) break;
        // raw input gradient r[l] (cin channels at this level)
        ConvArgs a{};
        a.im = im;
        a.pim = im;
        a.P = Pl;
        a.Pp = Pl;
        a.A = w.dz[l];
        a.B = nullptr;
        a.W = pk.bwd[l];
        a.kp = nn_kp_bwd(l);
        a.out = w.r[l];
        conv_input_grad(l, a, conv_grid, st);
        // route it to the producers of this convolution's input
        switch (l) {
            case 13:  // [e0, up(d1)]: d1's gradient from the upsampled part
                k_up_combine<<<cdiv(level_pixels(lv.lv[1]) * 48, 256), 256, 0, st>>>(lv.lv[0], lv.lv[1],
                    level_pixels(lv.lv[1]), 48, w.r[13], 80, 32, w.act[12], w.dz[12]);
                break;
            case 11:  // [e1, up(d2)]
                k_up_combine<<<cdiv(level_pixels(lv.lv[2]) * 64, 256), 256, 0, st>>>(lv.lv[1], lv.lv[2],
                    level_pixels(lv.lv[2]), 64, w.r[11], 112, 48, w.act[10], w.dz[10]);
                break;
            case 9:  // [e2, up(b)]
                k_up_combine<<<cdiv(level_pixels(lv.lv[3]) * 96, 256), 256, 0, st>>>(lv.lv[2], lv.lv[3],
                    level_pixels(lv.lv[3]), 96, w.r[9], 160, 64, w.act[8], w.dz[8]);
                break;
            case 6:  // pool(e2) -> e2, plus the skip into convolution 9
                k_skip_combine<<<cdiv(level_pixels(lv.lv[2]) * 64, 256), 256, 0, st>>>(lv.lv[2], lv.lv[3],
                    level_pixels(lv.lv[2]), 64, w.r[9], 160, w.r[6], w.act[5], w.dz[5]);
                break;
            case 4:  // pool(e1) -> e1, plus the skip into convolution 11
                k_skip_combine<<<cdiv(level_pixels(lv.lv[1]) * 48, 256), 256, 0, st>>>(lv.lv[1], lv.lv[2],
                    level_pixels(lv.lv[1]), 48, w.r[11], 112, w.r[4], w.act[3], w.dz[3]);
                break;
            case 2:  // pool(e0) -> e0, plus the skip into convolution 13
                k_skip_combine<<<cdiv(level_pixels(lv.lv[0]) * 32, 256), 256, 0, st>>>(lv.lv[0], lv.lv[1],
                    level_pixels(lv.lv[0]), 32, w.r[13], 80, w.r[2], w.act[1], w.dz[1]);
                break;
            default: {  // single input from convolution l - 1 at this level
                const size_t n = (size_t)Pl * cin;
                k_mask<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, w.r[l], w.act[l - 1], w.dz[l - 1]);
                break;
            }
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
