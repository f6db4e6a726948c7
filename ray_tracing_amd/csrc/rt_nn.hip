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

// Output tiles of the convolution kernel: kCvW pixels of ROWS rows of one
// image, one row per wave (four 16-pixel MFMA tiles); 8 rows at full size
// (fewer re-read halo rows), 4 below (less LDS per block: more blocks per
// CU for the wider layers).  The input rows y0 - 1 .. y0 + ROWS, pixels
// x0 - 1 .. x0 + kCvW, are staged in LDS; a decoder's coarser input is
// staged at its own resolution (the parents of those pixels: ROWS / 2 + 2
// rows of kCvCoarseRun).
constexpr int kCvW = 64;
constexpr int kCvRun = kCvW + 2;
constexpr int kCvCoarseRun = kCvW / 2 + 2;

// LDS row stride (bf16) of a staged pixel of c channels: at least c, and an
// odd number of 16-byte units, so that the 8 pixels one quarter-wave of
// 16-byte reads touches fall on distinct bank groups
__host__ __device__ constexpr int cv_row(int c) { return (c / 8) % 2 == 1 ? c : c + 8; }

// tiles of kCvW x rows pixels of a level: per image, rows of tiles
__device__ __forceinline__ void tile_locate(const NnImages& im, int rows, int b, int* img, int* x0, int* y0) {
    int i = 0;
    for (;;) {
        const int w = im.width[i], h = (im.offset[i + 1] - im.offset[i]) / w;
        const int tx = (w + kCvW - 1) / kCvW, n = tx * ((h + rows - 1) / rows);
        if (b < n || i + 1 >= im.count) {
            *img = i;
            *y0 = (b / tx) * rows;
            *x0 = (b - (b / tx) * tx) * kCvW;
            return;
        }
        b -= n;
        ++i;
    }
}

// One block computes out[tile][COUT] for a kCvW x ROWS tile as the
// transposed GEMM out^T = W X^T: the MFMA A operand is the weights (row =
// output channel, 8 consecutive K per lane, one 16-byte load from the
// L1-resident weight matrix, fetched one K step ahead), the B operand the
// input (column = pixel, 8 channels of one tap per lane), read from the
// staged tile with one 16-byte LDS load at the lane's own tap.  K runs over
// (tap, channel) pairs without per-tap padding (k = tap * CIN + channel,
// padded once to a multiple of 32; padding lanes read zero).  The fine tile
// holds zeros outside the image; a decoder's coarse channels are read at
// the parent of the tap's pixel and zeroed where that pixel is outside the
// image.  A lane ends with 4 consecutive output channels of one pixel and
// stores them as 8 bytes.  MODE 0: out = relu(acc + bias); MODE 1: out = acc.
template <int C0, int C1, int COUT, int MODE, int ROWS>
__global__ __launch_bounds__(64 * ROWS) void k_uconv(ConvArgs a) {
    constexpr int kCvThreads = 64 * ROWS;
    constexpr int kCvInRows = ROWS + 2;
    constexpr int kCvCoarseRows = ROWS / 2 + 2;
    constexpr int CIN = C0 + C1;
    constexpr int KP = nn_kp(CIN);
    constexpr int STEPS = KP / 32;
    constexpr int NT = COUT / 16;
    constexpr int MT = kCvW / 16;
    constexpr int RA = cv_row(C0);
    constexpr int RB = C1 > 0 ? cv_row(C1) : 8;
    static_assert(CIN % 8 == 0 && C0 % 8 == 0 && COUT % 16 == 0, "channel layout");
    __shared__ __align__(16) uint16_t sa[kCvInRows * kCvRun * RA];
    __shared__ __align__(16) uint16_t sb[C1 > 0 ? kCvCoarseRows * kCvCoarseRun * RB : 8];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int r16 = lane & 15, kq = lane >> 4;
    int img, x0, y0;
    tile_locate(a.im, ROWS, blockIdx.x, &img, &x0, &y0);
    const int w = a.im.width[img], h = (a.im.offset[img + 1] - a.im.offset[img]) / w;
    const int base = a.im.offset[img];
    for (int u = tid; u < kCvInRows * kCvRun * (C0 / 8); u += kCvThreads) {
        const int rest = u / (C0 / 8), c = (u - rest * (C0 / 8)) * 8;
        const int r = rest / kCvRun, q = rest - r * kCvRun;
        const int x = x0 - 1 + q, y = y0 - 1 + r;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (x >= 0 && x < w && y >= 0 && y < h)
            v = *reinterpret_cast<const uint4*>(a.A + (size_t)(base + y * w + x) * C0 + c);
        *reinterpret_cast<uint4*>(sa + rest * RA + c) = v;
    }
    const int px0 = (x0 - 1) >> 1, py0 = (y0 - 1) >> 1;
    if (C1 > 0) {
        const int pw = a.pim.width[img];
        const int ph = (a.pim.offset[img + 1] - a.pim.offset[img]) / pw;
        for (int u = tid; u < kCvCoarseRows * kCvCoarseRun * (C1 / 8); u += kCvThreads) {
            const int rest = u / (C1 / 8), c = (u - rest * (C1 / 8)) * 8;
            const int r = rest / kCvCoarseRun, q = rest - r * kCvCoarseRun;
            const int x = px0 + q, y = py0 + r;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (x >= 0 && x < pw && y >= 0 && y < ph)
                v = *reinterpret_cast<const uint4*>(a.B + (size_t)(a.pim.offset[img] + y * pw + x) * C1 + c);
            *reinterpret_cast<uint4*>(sb + rest * RB + c) = v;
        }
    }
    __syncthreads();
    f32x4 acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[m][nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const uint16_t* wrow = a.W + (size_t)r16 * KP + kq * 8;
    bf16x8 wv[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) wv[nt] = *reinterpret_cast<const bf16x8*>(wrow + (size_t)nt * 16 * KP);
    const bf16x8 zero8 = __builtin_bit_cast(bf16x8, make_uint4(0u, 0u, 0u, 0u));
    const int y = y0 + wave;
#pragma unroll 2
    for (int step = 0; step < STEPS; ++step) {
        // the next step's weights, fetched while this step's MFMAs run
        bf16x8 wn[NT];
        if (step + 1 < STEPS) {
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                wn[nt] = *reinterpret_cast<const bf16x8*>(wrow + (size_t)nt * 16 * KP + (step + 1) * 32);
        }
        // this lane's 8 K values: channels ch .. ch + 7 of tap t
        const int k0 = step * 32 + kq * 8;
        const bool kin = k0 < 9 * CIN;
        const int t = kin ? k0 / CIN : 0;
        const int ch = k0 - t * CIN;
        const int dy = t / 3 - 1, dx = t - (t / 3) * 3 - 1;
        // this lane's pixel in tile m is (x0 + 16 m + r16, y); the tap reads
        // fine slot (wave + dy + 1, 16 m + r16 + dx + 1)
        const uint16_t* xa = sa + ((wave + dy + 1) * kCvRun + r16 + dx + 1) * RA;
        bf16x8 xv[MT];
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            if (C1 == 0) {
                xv[m] = *reinterpret_cast<const bf16x8*>(xa + m * 16 * RA + ch);
            } else {
                const int xs = x0 + m * 16 + r16 + dx, ys = y + dy;
                const bool in = xs >= 0 && xs < w && ys >= 0 && ys < h;
                const uint16_t* pb = sb + (((ys >> 1) - py0) * kCvCoarseRun + ((xs >> 1) - px0)) * RB;
                const uint16_t* src = ch < C0 ? xa + m * 16 * RA + ch : pb + (ch - C0);
                xv[m] = *reinterpret_cast<const bf16x8*>(src);
                if (ch >= C0 && !in) xv[m] = zero8;
            }
            if (!kin) xv[m] = zero8;
        }
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                acc[m][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wv[nt], xv[m], acc[m][nt], 0, 0, 0);
        if (step + 1 < STEPS) {
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) wv[nt] = wn[nt];
        }
    }
    // D[i][j]: i = output channel 4 kq + e of the tile, j = pixel r16
    if (y >= h) return;
#pragma unroll
    for (int m = 0; m < MT; ++m) {
        const int x = x0 + m * 16 + r16;
        if (x >= w) continue;
        const size_t p = (size_t)(base + y * w + x);
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
            *reinterpret_cast<uint2*>(a.out + p * COUT + co) = packed;
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

// 8 bf16 values of one 16-byte load as floats, and back
__device__ __forceinline__ void unpack8(uint4 u, float f[8]) {
    const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        f[2 * k] = from_bf16((uint16_t)(w[k] & 0xFFFFu));
        f[2 * k + 1] = from_bf16((uint16_t)(w[k] >> 16));
    }
}
__device__ __forceinline__ uint4 pack8(const float f[8]) {
    uint32_t w[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) w[k] = (uint32_t)to_bf16(f[2 * k]) | ((uint32_t)to_bf16(f[2 * k + 1]) << 16);
    return make_uint4(w[0], w[1], w[2], w[3]);
}
// the bf16 words of v where the matching words of act are > 0, else 0
__device__ __forceinline__ uint4 relu_mask8(uint4 v, uint4 act) {
    const uint32_t a[4] = {act.x, act.y, act.z, act.w};
    uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t lo = from_bf16((uint16_t)(a[k] & 0xFFFFu)) > 0.0f ? 0x0000FFFFu : 0u;
        const uint32_t hi = from_bf16((uint16_t)(a[k] >> 16)) > 0.0f ? 0xFFFF0000u : 0u;
        w[k] &= lo | hi;
    }
    return make_uint4(w[0], w[1], w[2], w[3]);
}

// dz = r where act > 0; 8 channels per thread
__global__ __launch_bounds__(256) void k_mask(size_t n8, const uint16_t* r, const uint16_t* act,
    uint16_t* dz) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n8) return;
    reinterpret_cast<uint4*>(dz)[i] = relu_mask8(reinterpret_cast<const uint4*>(r)[i],
        reinterpret_cast<const uint4*>(act)[i]);
}

// dz at level l+1 = sum over the covered level-l pixels of r[:, off .. off + C)
// (the upsampled input's gradient), where act > 0; 8 channels per thread
__global__ __launch_bounds__(256) void k_up_combine(NnImages im, NnImages pim, int Pp, int C,
    const uint16_t* r, int cr, int off, const uint16_t* act, uint16_t* dz) {
    const int c8 = C / 8;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= Pp * c8) return;
    const int pp = i / c8, c = (i - pp * c8) * 8;
    const Pix q = locate(pim, pp);
    const int w = im.width[q.img], h = image_height(im, q.img), base = im.offset[q.img];
    float s[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    for (int dy = 0; dy < 2; ++dy) {
        for (int dx = 0; dx < 2; ++dx) {
            const int x = 2 * q.x + dx, y = 2 * q.y + dy;
            if (x < w && y < h) {
                float f[8];
                unpack8(*reinterpret_cast<const uint4*>(r + (size_t)(base + y * w + x) * cr + off + c), f);
#pragma unroll
                for (int k = 0; k < 8; ++k) s[k] += f[k];
            }
        }
    }
    const size_t o = (size_t)pp * C + c;
    *reinterpret_cast<uint4*>(dz + o) = relu_mask8(pack8(s), *reinterpret_cast<const uint4*>(act + o));
}

// dz at level l = (skip gradient rs[:, 0..C) + pooled-path gradient rp at the
// parent / pixels the parent covers), where act > 0; 8 channels per thread
__global__ __launch_bounds__(256) void k_skip_combine(NnImages im, NnImages pim, int P, int C,
    const uint16_t* rs, int cs, const uint16_t* rp, const uint16_t* act, uint16_t* dz) {
    const int c8 = C / 8;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= P * c8) return;
    const int p = i / c8, c = (i - p * c8) * 8;
    const Pix q = locate(im, p);
    const int px = q.x >> 1, py = q.y >> 1;
    const int n = min(2, q.w - 2 * px) * min(2, q.h - 2 * py);
    const int pp = pim.offset[q.img] + py * pim.width[q.img] + px;
    float a[8], b[8];
    unpack8(*reinterpret_cast<const uint4*>(rs + (size_t)p * cs + c), a);
    unpack8(*reinterpret_cast<const uint4*>(rp + (size_t)pp * C + c), b);
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] = a[k] + b[k] / (float)n;
    const size_t o = (size_t)p * C + c;
    *reinterpret_cast<uint4*>(dz + o) = relu_mask8(pack8(a), *reinterpret_cast<const uint4*>(act + o));
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
    __shared__ float red[4 * (kNnHeadParams + 1)];
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
    // per-block sums of the head gradient terms and the loss: a butterfly
    // within each wave, then the four wave sums in wave order (fixed order)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int i = 0; i < kNnHeadParams + 1; ++i) {
        float v = 0.0f;
        if (valid) {
            if (i < 3 * kNnOut) v = dout[i / kNnOut] * h[i % kNnOut];
            else if (i < kNnHeadParams) v = dout[i - 3 * kNnOut];
            else v = loss;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0) red[wave * (kNnHeadParams + 1) + i] = v;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kNnHeadParams + 1; i += 256) {
        const int n = kNnHeadParams + 1;
        partial[(size_t)blockIdx.x * n + i] = ((red[i] + red[n + i]) + red[2 * n + i]) + red[3 * n + i];
    }
}

struct WgradArgs {
    NnImages im, pim;
    int segs;             // row segments of kWgSeg pixels at this level
    int chunk_segs;       // segments per block
    const uint16_t* dz;   // [P][cout]
    const uint16_t* A;    // [P][C0]
    const uint16_t* B;    // [P coarser][C1]
    float* partial;       // [chunks][cout][nw16]
};

constexpr int kWgSeg = 64;   // pixels per stage: one segment of one image row (two MFMA K steps)

// LDS row stride (bf16 elements) of a [pixel][channel] tile: at least c, and
// 2 * stride = 32 mod 128 bytes, so that the four rows one 16-lane
// transposed read touches fall on distinct banks
__host__ __device__ constexpr int wg_row(int c) { return c + ((16 - c % 64) + 64) % 64; }

typedef short s16x4 __attribute__((ext_vector_type(4)));

// ds_read_b64_tr_b16: within each group of 16 lanes, lane j reads 4 bf16 at
// its address; lane i then holds element (i % 4) of lanes 4 k + i / 4,
// k = 0..3.  With lane j addressing row r0 + j / 4, column c0 + 4 (j % 4) of a
// row-major [row][column] tile, lane i holds column c0 + i of rows r0..r0+3:
// a K-contiguous MFMA operand read straight from a pixel-major tile.  Two
// reads, 4 rows apart, give the 8 rows of a 16x16x32 operand.
__device__ __forceinline__ bf16x8 tr_operand(const uint16_t* tile, int row_stride, int r0, int c0,
    int j16) {
    const uint16_t* p = tile + (r0 + (j16 >> 2)) * row_stride + c0 + 4 * (j16 & 3);
    const s16x4 lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16((__attribute__((address_space(3))) s16x4*)p);
    const s16x4 hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16(
        (__attribute__((address_space(3))) s16x4*)(p + 4 * row_stride));
    return __builtin_bit_cast(bf16x8, __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7));
}

struct Seg {
    int img, base, w, h, y, x0;
};

// segment s: image, row and first pixel (segments run along each row, rows
// in order, images back to back)
__device__ __forceinline__ Seg seg_locate(const NnImages& im, int s) {
    int i = 0;
    Seg r;
    for (;;) {
        r.w = im.width[i];
        r.h = (im.offset[i + 1] - im.offset[i]) / r.w;
        const int n = r.h * ((r.w + kWgSeg - 1) / kWgSeg);
        if (s < n || i + 1 >= im.count) break;
        s -= n;
        ++i;
    }
    const int per_row = (r.w + kWgSeg - 1) / kWgSeg;
    r.img = i;
    r.base = im.offset[i];
    r.y = s / per_row;
    r.x0 = (s - r.y * per_row) * kWgSeg;
    return r;
}

// Weight gradient of one convolution as GEMMs with K = pixels:
//   G[co][t CIN + ci] = sum_p dz[p][co] X[p + off(t)][ci]   (taps t)
//   G[co][9 CIN]      = sum_p dz[p][co]                      (the bias)
// Block chunk * (9 / TPB) + g covers the segments of one chunk and the taps
// g TPB .. g TPB + TPB - 1.  Per segment it stages dz and the input rows the
// taps read (pixels x0 - 1 .. x0 + kWgSeg, zero outside the image) in LDS as
// [pixel][channel] with 16-byte stores; both MFMA operands (8 consecutive
// pixels of one channel per lane) come from transposed LDS reads.  Wave w
// owns the (tap, 16-channel input) tiles w, w + 4, ... of the block with all
// output tiles, so each operand fragment it reads feeds several MFMAs.
// Partial sums per chunk are reduced in chunk order afterwards
// (deterministic).
template <int C0, int C1, int COUT, int TPB>
__global__ __launch_bounds__(256) void k_wgrad(WgradArgs a) {
    constexpr int CIN = C0 + C1;
    constexpr int MT = COUT / 16, CT = CIN / 16;
    constexpr int TN = TPB * CT;
    // the four waves as a WN x WM grid over the (tap, input) x output tiles
    constexpr int WN = TN >= 4 ? 4 : (TN >= 2 ? 2 : 1);
    constexpr int WM = 4 / WN;
    constexpr int NPW = (TN + WN - 1) / WN, MPW = (MT + WM - 1) / WM;
    constexpr int ROWS = TPB == 9 ? 3 : 1;
    constexpr int GROUPS = 9 / TPB;
    constexpr int RZ = wg_row(COUT), RX = wg_row(CIN);
    constexpr int RUN = kWgSeg + 2;
    constexpr int NWC = ((9 * CIN) / 16 + 1) * 16;
    static_assert(CIN % 16 == 0 && C0 % 8 == 0 && COUT % 16 == 0 && 9 % TPB == 0, "layout");
    __shared__ __align__(16) uint16_t sdz[kWgSeg * RZ];
    __shared__ __align__(16) uint16_t sx[ROWS * RUN * RX];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wn = wave % WN, wm = wave / WN;
    const int j16 = lane & 15, g4 = lane >> 4;
    const int chunk = (int)blockIdx.x / GROUPS, grp = (int)blockIdx.x - chunk * GROUPS;
    const int t0 = grp * TPB;
    const bool bias = grp == 0 && wn == 0;   // the wn = 0 waves cover every output tile once
    f32x4 acc[MPW][NPW];
    f32x4 accb[MPW];
#pragma unroll
    for (int m = 0; m < MPW; ++m) {
        accb[m] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int n = 0; n < NPW; ++n) acc[m][n] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    }
    const bf16x8 ones = __builtin_bit_cast(bf16x8, make_uint4(0x3F803F80u, 0x3F803F80u, 0x3F803F80u, 0x3F803F80u));
    const int s_begin = chunk * a.chunk_segs;
    const int s_end = min(a.segs, s_begin + a.chunk_segs);
    // software pipeline: the 16-byte pieces of segment s + 1 are fetched into
    // registers while the MFMAs of segment s run
    constexpr int NZ = (kWgSeg * (COUT / 8) + 255) / 256;
    constexpr int NX = (ROWS * RUN * (CIN / 8) + 255) / 256;
    uint4 rz[NZ], rx[NX];
    auto fetch = [&](int seg) {
        const Seg sg = seg_locate(a.im, seg);
#pragma unroll
        for (int k = 0; k < NZ; ++k) {
            const int u = tid + 256 * k;
            const int q = u / (COUT / 8), c = (u - q * (COUT / 8)) * 8;
            rz[k] = make_uint4(0u, 0u, 0u, 0u);
            if (u < kWgSeg * (COUT / 8) && sg.x0 + q < sg.w)
                rz[k] = *reinterpret_cast<const uint4*>(a.dz + (size_t)(sg.base + sg.y * sg.w + sg.x0 + q) * COUT + c);
        }
#pragma unroll
        for (int k = 0; k < NX; ++k) {
            const int u = tid + 256 * k;
            const int rest = u / (CIN / 8), c = (u - rest * (CIN / 8)) * 8;
            const int r = rest / RUN, q = rest - r * RUN;
            const int dy = ROWS == 3 ? r - 1 : t0 / 3 - 1;
            const int x = sg.x0 - 1 + q, y = sg.y + dy;
            rx[k] = make_uint4(0u, 0u, 0u, 0u);
            if (u < ROWS * RUN * (CIN / 8) && x >= 0 && x < sg.w && y >= 0 && y < sg.h) {
                if (C1 == 0 || c < C0)
                    rx[k] = *reinterpret_cast<const uint4*>(a.A + (size_t)(sg.base + y * sg.w + x) * C0 + c);
                else
                    rx[k] = *reinterpret_cast<const uint4*>(a.B + (size_t)parent_index(a.pim, sg.img, x, y) * C1 + (c - C0));
            }
        }
    };
    if (s_begin < s_end) fetch(s_begin);
    for (int s = s_begin; s < s_end; ++s) {
#pragma unroll
        for (int k = 0; k < NZ; ++k) {
            const int u = tid + 256 * k;
            const int q = u / (COUT / 8), c = (u - q * (COUT / 8)) * 8;
            if (u < kWgSeg * (COUT / 8)) *reinterpret_cast<uint4*>(sdz + q * RZ + c) = rz[k];
        }
#pragma unroll
        for (int k = 0; k < NX; ++k) {
            const int u = tid + 256 * k;
            const int rest = u / (CIN / 8), c = (u - rest * (CIN / 8)) * 8;
            if (u < ROWS * RUN * (CIN / 8)) *reinterpret_cast<uint4*>(sx + rest * RX + c) = rx[k];
        }
        __syncthreads();
        if (s + 1 < s_end) fetch(s + 1);
#pragma unroll
        for (int ks = 0; ks < kWgSeg / 32; ++ks) {
            const int r0 = ks * 32 + g4 * 8;
            bf16x8 af[MPW];
#pragma unroll
            for (int m = 0; m < MPW; ++m) {
                const int mt = wm + WM * m;
                if (mt < MT) af[m] = tr_operand(sdz, RZ, r0, mt * 16, j16);
            }
            // B fragments in batches of four (a tile past the last reads the
            // last one and is not used), so the reads of a batch overlap and
            // its MFMAs follow without waiting on one read at a time
            constexpr int NB = NPW < 4 ? NPW : 4;
#pragma unroll
            for (int n0 = 0; n0 < NPW; n0 += NB) {
                bf16x8 bfr[NB];
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    const int nt = min(wn + WN * min(n0 + b, NPW - 1), TN - 1);
                    const int t = t0 + nt / CT, ct = nt - (nt / CT) * CT;
                    const int dx = t % 3 - 1;
                    const int row = ROWS == 3 ? t / 3 : 0;
                    bfr[b] = tr_operand(sx + row * RUN * RX, RX, r0 + dx + 1, ct * 16, j16);
                }
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    const int n = n0 + b;
                    if (n < NPW && wn + WN * n < TN) {
#pragma unroll
                        for (int m = 0; m < MPW; ++m)
                            if (wm + WM * m < MT)
                                acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[m], bfr[b], acc[m][n], 0, 0, 0);
                    }
                }
            }
            if (bias) {
#pragma unroll
                for (int m = 0; m < MPW; ++m)
                    if (wm + WM * m < MT)
                        accb[m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[m], ones, accb[m], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // D[i][j] of a tile: i = output channel 4 g4 + e, j = input channel j16
    float* dst = a.partial + (size_t)chunk * COUT * NWC;
#pragma unroll
    for (int m = 0; m < MPW; ++m) {
        const int mt = wm + WM * m;
        if (mt >= MT) continue;
#pragma unroll
        for (int n = 0; n < NPW; ++n) {
            const int nt = wn + WN * n;
            if (nt >= TN) continue;
            const int t = t0 + nt / CT, ct = nt - (nt / CT) * CT;
#pragma unroll
            for (int e = 0; e < 4; ++e)
                dst[(size_t)(mt * 16 + g4 * 4 + e) * NWC + t * CIN + ct * 16 + j16] = acc[m][n][e];
        }
        if (bias && j16 == 0) {
#pragma unroll
            for (int e = 0; e < 4; ++e) dst[(size_t)(mt * 16 + g4 * 4 + e) * NWC + 9 * CIN] = accb[m][e];
        }
    }
}

// sum the chunks' partials of one convolution into the gradient vector
// (W[co][tap][ci] then b[co]): block of 16 outputs x 16 chunk lanes, lane c
// summing chunks c, c + 16, ... in order, then the 16 lane sums in order
// (a fixed order: deterministic)
__global__ __launch_bounds__(256) void k_uwgrad_reduce(int chunks, int cout, int cin,
    const float* partial, float* grad) {
    __shared__ float part[16][17];
    const int o = threadIdx.x & 15, c = threadIdx.x >> 4;
    const int i = blockIdx.x * 16 + o;
    const int total = cout * (9 * cin + 1);
    const int nw16 = ((9 * cin) / 16 + 1) * 16;
    float s = 0.0f;
    if (i < total) {
        int co, n;
        if (i < cout * 9 * cin) {
            co = i / (9 * cin);
            n = i - co * 9 * cin;
        } else {
            co = i - cout * 9 * cin;
            n = 9 * cin;
        }
        for (int b = c; b < chunks; b += 16) s += partial[((size_t)b * cout + co) * nw16 + n];
    }
    part[c][o] = s;
    __syncthreads();
    if (c == 0 && i < total) {
        float t = 0.0f;
        for (int k = 0; k < 16; ++k) t += part[k][o];
        grad[i] = t;
    }
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
        const int co = i / kpf, k = i - co * kpf;
        const int t = k / cin, ci = k - t * cin;
        fwd[i] = k < 9 * cin ? to_bf16(W[((size_t)co * 9 + t) * cin + ci]) : (uint16_t)0;
    }
    if (bwd && i < cin * kpb) {
        const int ci = i / kpb, k = i - ci * kpb;
        const int tp = k / cout, co = k - tp * cout;
        bwd[i] = k < 9 * cout ? to_bf16(W[((size_t)co * 9 + (8 - tp)) * cin + ci]) : (uint16_t)0;
    }
    if (i < cout) bias[i] = W[(size_t)cout * 9 * cin + i];
}

__global__ void k_copy(int n, const float* src, float* dst) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

inline int cdiv(int a, int b) { return (a + b - 1) / b; }

inline int level_pixels(const NnImages& im) { return im.offset[im.count]; }

// tiles of kCvW x rows pixels at a level
int level_tiles(const NnImages& im, int rows) {
    int n = 0;
    for (int i = 0; i < im.count; ++i) {
        const int w = im.width[i], h = (im.offset[i + 1] - im.offset[i]) / w;
        n += cdiv(w, kCvW) * cdiv(h, rows);
    }
    return n;
}

// forward (MODE 0) or input-gradient (MODE 1) launch of a convolution: one
// block per output tile of `rows` rows (4 or 8)
template <int C0, int C1, int COUT, int MODE>
void conv_launch(const ConvArgs& a, int rows, hipStream_t st) {
    const int tiles = level_tiles(a.im, rows);
    if (tiles <= 0) return;
    if (rows == 8) k_uconv<C0, C1, COUT, MODE, 8><<<tiles, 512, 0, st>>>(a);
    else k_uconv<C0, C1, COUT, MODE, 4><<<tiles, 256, 0, st>>>(a);
}

}  // namespace

// the tiling's rows at a level (0: 8 when the widest image is >= 160 pixels)
int nn_tile_rows(const NnTiling& t, const NnImages& im, int level) {
    if (t.rows[level] == 4 || t.rows[level] == 8) return t.rows[level];
    int wmax = 0;
    for (int i = 0; i < im.count; ++i) wmax = std::max(wmax, im.width[i]);
    return wmax >= 160 ? 8 : 4;
}

namespace {

void conv_forward(int l, const ConvArgs& a, int rows, hipStream_t st) {
    switch (l) {
        case 0: conv_launch<16, 0, 32, 0>(a, rows, st); break;
        case 1: conv_launch<32, 0, 32, 0>(a, rows, st); break;
        case 2: conv_launch<32, 0, 48, 0>(a, rows, st); break;
        case 3: conv_launch<48, 0, 48, 0>(a, rows, st); break;
        case 4: conv_launch<48, 0, 64, 0>(a, rows, st); break;
        case 5: conv_launch<64, 0, 64, 0>(a, rows, st); break;
        case 6: conv_launch<64, 0, 96, 0>(a, rows, st); break;
        case 7: conv_launch<96, 0, 96, 0>(a, rows, st); break;
        case 8: conv_launch<96, 0, 96, 0>(a, rows, st); break;
        case 9: conv_launch<64, 96, 64, 0>(a, rows, st); break;
        case 10: conv_launch<64, 0, 64, 0>(a, rows, st); break;
        case 11: conv_launch<48, 64, 48, 0>(a, rows, st); break;
        case 12: conv_launch<48, 0, 48, 0>(a, rows, st); break;
        case 13: conv_launch<32, 48, 32, 0>(a, rows, st); break;
        default: conv_launch<32, 0, 32, 0>(a, rows, st); break;
    }
}

// input gradient of convolution l: the output gradient (cout channels) through
// the transposed weights into cin raw channels
void conv_input_grad(int l, const ConvArgs& a, int rows, hipStream_t st) {
    switch (l) {
        case 1: conv_launch<32, 0, 32, 1>(a, rows, st); break;
        case 2: conv_launch<48, 0, 32, 1>(a, rows, st); break;
        case 3: conv_launch<48, 0, 48, 1>(a, rows, st); break;
        case 4: conv_launch<64, 0, 48, 1>(a, rows, st); break;
        case 5: conv_launch<64, 0, 64, 1>(a, rows, st); break;
        case 6: conv_launch<96, 0, 64, 1>(a, rows, st); break;
        case 7: conv_launch<96, 0, 96, 1>(a, rows, st); break;
        case 8: conv_launch<96, 0, 96, 1>(a, rows, st); break;
        case 9: conv_launch<64, 0, 160, 1>(a, rows, st); break;
        case 10: conv_launch<64, 0, 64, 1>(a, rows, st); break;
        case 11: conv_launch<48, 0, 112, 1>(a, rows, st); break;
        case 12: conv_launch<48, 0, 48, 1>(a, rows, st); break;
        case 13: conv_launch<32, 0, 80, 1>(a, rows, st); break;
        default: conv_launch<32, 0, 32, 1>(a, rows, st); break;
    }
}

// row segments of kWgSeg pixels at a level
int level_segments(const NnImages& im) {
    int n = 0;
    for (int i = 0; i < im.count; ++i) {
        const int w = im.width[i], h = (im.offset[i + 1] - im.offset[i]) / w;
        n += h * cdiv(w, kWgSeg);
    }
    return n;
}

// taps per weight-gradient block: all nine at full size with at most 32
// input channels (most pixels, fewest channels: the most MFMAs per staged
// byte), otherwise a row of three at the two finer levels and one below
// (more blocks, fewer accumulators per wave: at most 16 tiles per block)
constexpr int wgrad_tpb(int l) {
    return nn_conv(l).level == 0 ? (nn_cin(l) <= 32 ? 9 : 3) : (nn_conv(l).level == 1 ? 3 : 1);
}

// segments per weight-gradient block: the partial sums (chunks x weights)
// are written and read once more, so chunks are bounded: 384 at full size,
// 192 at half size, 64 below (384 to 576 blocks per convolution)
int wgrad_chunk_segs(int l, int segs) {
    const int level = nn_conv(l).level;
    const int chunks = std::max(1, std::min(segs, level == 0 ? 384 : (level == 1 ? 192 : 64)));
    return std::max(1, cdiv(segs, chunks));
}

template <int L>
void wgrad_launch(const WgradArgs& a, int chunks, hipStream_t st) {
    constexpr NnConv c = nn_conv(L);
    constexpr int TPB = wgrad_tpb(L);
    if (chunks > 0) k_wgrad<c.c0, c.c1, c.cout, TPB><<<chunks * (9 / TPB), 256, 0, st>>>(a);
}

void conv_wgrad(int l, const WgradArgs& a, int chunks, hipStream_t st) {
    switch (l) {
        case 0: wgrad_launch<0>(a, chunks, st); break;
        case 1: wgrad_launch<1>(a, chunks, st); break;
        case 2: wgrad_launch<2>(a, chunks, st); break;
        case 3: wgrad_launch<3>(a, chunks, st); break;
        case 4: wgrad_launch<4>(a, chunks, st); break;
        case 5: wgrad_launch<5>(a, chunks, st); break;
        case 6: wgrad_launch<6>(a, chunks, st); break;
        case 7: wgrad_launch<7>(a, chunks, st); break;
        case 8: wgrad_launch<8>(a, chunks, st); break;
        case 9: wgrad_launch<9>(a, chunks, st); break;
        case 10: wgrad_launch<10>(a, chunks, st); break;
        case 11: wgrad_launch<11>(a, chunks, st); break;
        case 12: wgrad_launch<12>(a, chunks, st); break;
        case 13: wgrad_launch<13>(a, chunks, st); break;
        default: wgrad_launch<14>(a, chunks, st); break;
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
        const int segs = level_segments(lv.lv[c.level]);
        const int chunks = cdiv(segs, wgrad_chunk_segs(l, segs));
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
    const float4* target, float4* out, const NnTiling& tiling, hipStream_t st) {
    const int P = level_pixels(lv.lv[0]);
    if (P <= 0) return hipSuccess;
    k_features<<<cdiv(P, 256), 256, 0, st>>>(P, beauty, normal, albedo, depth, w.feat);
    for (int l = 0; l < kNnConvs; ++l) {
        ConvArgs a = conv_args(lv, l, w);
        a.W = pk.fwd[l];
        a.kp = nn_kp_fwd(l);
        a.bias = pk.bias[l];
        a.out = w.act[l];
        conv_forward(l, a, nn_tile_rows(tiling, a.im, nn_conv(l).level), st);
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
    float* loss, const NnTiling& tiling, hipStream_t st) {
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
        g.segs = level_segments(im);
        g.chunk_segs = wgrad_chunk_segs(l, g.segs);
        g.dz = w.dz[l];
        conv_inputs(l, w, &g.A, &g.B);
        g.partial = w.grad_partial;
        const int chunks = cdiv(g.segs, g.chunk_segs);
        conv_wgrad(l, g, chunks, st);
        k_uwgrad_reduce<<<cdiv(c.cout * (9 * cin + 1), 16), 256, 0, st>>>(
            chunks, c.cout, cin, w.grad_partial, grad + nn_layer_offset(l));
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
        conv_input_grad(l, a, nn_tile_rows(tiling, im, c.level), st);
        // route it to the producers of this convolution's input
        switch (l) {
            case 13:  // [e0, up(d1)]: d1's gradient from the upsampled part
                k_up_combine<<<cdiv(level_pixels(lv.lv[1]) * 48 / 8, 256), 256, 0, st>>>(lv.lv[0], lv.lv[1],
                    level_pixels(lv.lv[1]), 48, w.r[13], 80, 32, w.act[12], w.dz[12]);
                break;
            case 11:  // [e1, up(d2)]
                k_up_combine<<<cdiv(level_pixels(lv.lv[2]) * 64 / 8, 256), 256, 0, st>>>(lv.lv[1], lv.lv[2],
                    level_pixels(lv.lv[2]), 64, w.r[11], 112, 48, w.act[10], w.dz[10]);
                break;
            case 9:  // [e2, up(b)]
                k_up_combine<<<cdiv(level_pixels(lv.lv[3]) * 96 / 8, 256), 256, 0, st>>>(lv.lv[2], lv.lv[3],
                    level_pixels(lv.lv[3]), 96, w.r[9], 160, 64, w.act[8], w.dz[8]);
                break;
            case 6:  // pool(e2) -> e2, plus the skip into convolution 9
                k_skip_combine<<<cdiv(level_pixels(lv.lv[2]) * 64 / 8, 256), 256, 0, st>>>(lv.lv[2], lv.lv[3],
                    level_pixels(lv.lv[2]), 64, w.r[9], 160, w.r[6], w.act[5], w.dz[5]);
                break;
            case 4:  // pool(e1) -> e1, plus the skip into convolution 11
                k_skip_combine<<<cdiv(level_pixels(lv.lv[1]) * 48 / 8, 256), 256, 0, st>>>(lv.lv[1], lv.lv[2],
                    level_pixels(lv.lv[1]), 48, w.r[11], 112, w.r[4], w.act[3], w.dz[3]);
                break;
            case 2:  // pool(e0) -> e0, plus the skip into convolution 13
                k_skip_combine<<<cdiv(level_pixels(lv.lv[0]) * 32 / 8, 256), 256, 0, st>>>(lv.lv[0], lv.lv[1],
                    level_pixels(lv.lv[0]), 32, w.r[13], 80, w.r[2], w.act[1], w.dz[1]);
                break;
            default: {  // single input from convolution l - 1 at this level
                const size_t n8 = (size_t)Pl * cin / 8;
                k_mask<<<(unsigned)((n8 + 255) / 256), 256, 0, st>>>(n8, w.r[l], w.act[l - 1], w.dz[l - 1]);
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
