// rt_post.hip -- per-frame post passes over all cameras' pixels at once:
// temporal reprojection / accumulation and the edge-aware denoise filter,
// and the viewer's tone map to RGBA8.
//
// Pixels of all cameras lie back to back (CameraView::pixel_offset); one
// launch covers the whole rig.  All passes are bandwidth-bound.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "rt_camera.h"
#include "rt_kernels.h"
#include "rt_math.h"

namespace rtk {
namespace {

struct PixelId {
    int camera, x, y, local, width, height;
};

__device__ __forceinline__ PixelId pixel_id(const CameraView& cams, int p) {
    int c = 0;
    while (c + 1 < cams.count && p >= cams.pixel_offset[c + 1]) ++c;
    PixelId id;
    id.camera = c;
    id.local = p - cams.pixel_offset[c];
    id.width = cams.width[c];
    id.height = (cams.pixel_offset[c + 1] - cams.pixel_offset[c]) / id.width;
    id.y = id.local / id.width;
    id.x = id.local - id.y * id.width;
    return id;
}

__device__ __forceinline__ f3 decode_normal(float4 e) {
    return mk3(e.x * 2.0f - 1.0f, e.y * 2.0f - 1.0f, e.z * 2.0f - 1.0f);
}

// resolve_reprojection_kernel (src/realtime/gpu/programs.cu:1955-2067):
// reproject the pixel centre's surface into the previous camera, accept the
// history texel when normals agree (cos > 0.95) and depth matches within
// max(0.02, 5 %), then blend 1/(n+1) toward the current value.  Beauty is
// not blended when ReSTIR DI is on.  Writes flow and its trust mask.
__global__ __launch_bounds__(256) void k_resolve(CameraView cams, ResolveParams rp, float4* beauty,
    float4* normal, float* depth, const float4* hbeauty, const float4* hnormal,
    const float* hdepth, float2* flow, float* trust, int total) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= total) return;
    const PixelId id = pixel_id(cams, p);
    flow[p] = make_float2(0.0f, 0.0f);
    trust[p] = 0.0f;
    const int hl = rp.dyn->history_length[id.camera];
    if (hl <= 0 || rp.dyn->prev_valid[id.camera] == 0) return;
    const rt_camera& cam = cams.cams[id.camera];
    const rt_camera& pc = cams.cams[4 + id.camera];
    const float pxf = (float)id.x + 0.5f, pyf = (float)id.y + 0.5f;
    f3 o, d;
    camera_ray(cam, (double)pxf, (double)pyf, &o, &d);
    const float cd = depth[p];
    const f3 wp = o + d * cd;
    const f3 off = mk3(wp.x - (float)pc.origin[0], wp.y - (float)pc.origin[1], wp.z - (float)pc.origin[2]);
    const f3 bx = mk3((float)pc.basis_x[0], (float)pc.basis_x[1], (float)pc.basis_x[2]);
    const f3 by = mk3((float)pc.basis_y[0], (float)pc.basis_y[1], (float)pc.basis_y[2]);
    const f3 bz = mk3((float)pc.basis_z[0], (float)pc.basis_z[1], (float)pc.basis_z[2]);
    const f3 pcs = mk3(dot3(off, bx), dot3(off, by), dot3(off, bz));
    float ppx, ppy;
    project_camera(pc, pcs, &ppx, &ppy);
    const bool finite = isfinite(ppx) && isfinite(ppy);
    if (finite) flow[p] = make_float2(pxf - ppx, pyf - ppy);
    const int qx = finite ? (int)floorf(ppx) : -1;
    const int qy = finite ? (int)floorf(ppy) : -1;
    if (!finite || qx < 0 || qx >= id.width || qy < 0 || qy >= id.height) return;
    const int q = cams.pixel_offset[id.camera] + qy * id.width + qx;
    const float4 hn4 = hnormal[q];
    const float4 cn4 = normal[p];
    const float ndot = dot3(decode_normal(cn4), decode_normal(hn4));
    const float hd = hdepth[q];
    const float expected = len3(pcs);
    const float tol = fmaxf(0.02f, 0.05f * fmaxf(expected, hd));
    if (!(ndot > 0.95f && fabsf(expected - hd) < tol)) return;
    trust[p] = 1.0f;
    const float blend = 1.0f / (float)(hl + 1);
    if (rp.restir_enabled == 0) {
        const float4 hb = hbeauty[q];
        const float4 cb = beauty[p];
        const float cx = fminf(fmaxf(cb.x, 0.0f), 10.0f);
        const float cy = fminf(fmaxf(cb.y, 0.0f), 10.0f);
        const float cz = fminf(fmaxf(cb.z, 0.0f), 10.0f);
        beauty[p] = make_float4(hb.x + (cx - hb.x) * blend, hb.y + (cy - hb.y) * blend,
            hb.z + (cz - hb.z) * blend, 1.0f);
    }
    normal[p] = make_float4(hn4.x + (cn4.x - hn4.x) * blend, hn4.y + (cn4.y - hn4.y) * blend,
        hn4.z + (cn4.z - hn4.z) * blend, hn4.w + (cn4.w - hn4.w) * blend);
    depth[p] = hd + (cd - hd) * blend;
}

// One pass of the edge-aware a-trous filter: 5x5 B3-spline taps at spacing
// `step`, weighted by normal (cos^16), relative depth, albedo and relative
// luminance edge stops built from +, *, / only.  Pixels without a surface
// (depth 0) pass through and never contribute.
__constant__ float kB3[3] = {0.375f, 0.25f, 0.0625f};

__global__ __launch_bounds__(256) void k_denoise_pass(CameraView cams, int step, const float4* in,
    const float4* normal, const float4* albedo, const float* depth, float4* out, int total,
    const int32_t* gate) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= total) return;
    const PixelId id = pixel_id(cams, p);
    if (gate != nullptr && gate[id.camera] == 0) return;
    const float zp = depth[p];
    const float4 c4 = in[p];
    if (!(zp > 0.0f)) {
        out[p] = c4;
        return;
    }
    const int base = cams.pixel_offset[id.camera];
    const f3 cp = xyz(c4);
    const f3 np = decode_normal(normal[p]);
    const f3 ap = xyz(albedo[p]);
    const float lp = 0.2126f * cp.x + 0.7152f * cp.y + 0.0722f * cp.z;
    const float zs = 0.1f * (float)step * zp;
    f3 sum = mk3(0.0f, 0.0f, 0.0f);
    float wsum = 0.0f;
    for (int dy = -2; dy <= 2; ++dy) {
        const int qy = id.y + dy * step;
        if (qy < 0 || qy >= id.height) continue;
        for (int dx = -2; dx <= 2; ++dx) {
            const int qx = id.x + dx * step;
            if (qx < 0 || qx >= id.width) continue;
            const int q = base + qy * id.width + qx;
            const float zq = depth[q];
            if (!(zq > 0.0f)) continue;
            const f3 cq = xyz(in[q]);
            const f3 nq = decode_normal(normal[q]);
            const f3 aq = xyz(albedo[q]);
            const float lq = 0.2126f * cq.x + 0.7152f * cq.y + 0.0722f * cq.z;
            float wn = fmaxf(dot3(np, nq), 0.0f);
            wn = wn * wn;
            wn = wn * wn;
            wn = wn * wn;
            wn = wn * wn;
            const float ez = fabsf(zp - zq) / zs;
            const float wz = 1.0f / (1.0f + ez * ez);
            const float ea = (fabsf(ap.x - aq.x) + fabsf(ap.y - aq.y) + fabsf(ap.z - aq.z)) / 0.1f;
            const float wa = 1.0f / (1.0f + ea * ea);
            const float el = fabsf(lp - lq) / (4.0f * fmaxf(lp, lq) + 1e-3f);
            const float wl = 1.0f / (1.0f + el * el);
            const float k = kB3[dx < 0 ? -dx : dx] * kB3[dy < 0 ? -dy : dy];
            const float wt = k * wn * wz * wa * wl;
            sum = sum + cq * wt;
            wsum += wt;
        }
    }
    if (wsum > 0.0f) {
        const float inv = 1.0f / wsum;
        out[p] = make_float4(sum.x * inv, sum.y * inv, sum.z * inv, 1.0f);
    } else {
        out[p] = make_float4(cp.x, cp.y, cp.z, 1.0f);
    }
}

// Guard statistics of the learned denoiser, one block row per camera
// (blockIdx.y): over the surface pixels, non-finite outputs, the sums of
// log(1 + luminance) of input and output (a mean that single bright samples
// of a 1-spp input do not dominate), and their high-frequency energy
// (absolute luminance steps to the right-hand neighbour).  Block sums in LDS, one
// atomic per quantity and block.
__global__ __launch_bounds__(256) void k_nn_guard_stats(CameraView cams, const float4* in,
    const float4* out, const float* depth, NnGuard* guard) {
    const int c = blockIdx.y;
    const int base = cams.pixel_offset[c];
    const int n = cams.pixel_offset[c + 1] - base;
    const int w = cams.width[c];
    float v[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int p = base + i;
        if (!(depth[p] > 0.0f)) continue;
        const float4 a = in[p], b = out[p];
        const bool finite = isfinite(b.x) && isfinite(b.y) && isfinite(b.z);
        const float la = 0.2126f * a.x + 0.7152f * a.y + 0.0722f * a.z;
        const float lb = finite ? 0.2126f * b.x + 0.7152f * b.y + 0.0722f * b.z : 0.0f;
        v[0] += finite ? 0.0f : 1.0f;
        v[1] += log1pf(fmaxf(la, 0.0f));
        v[2] += log1pf(fmaxf(lb, 0.0f));
        v[5] += 1.0f;
        const int x = i % w;
        if (x + 1 < w && depth[p + 1] > 0.0f) {
            const float4 a1 = in[p + 1], b1 = out[p + 1];
            const float la1 = 0.2126f * a1.x + 0.7152f * a1.y + 0.0722f * a1.z;
            v[3] += fabsf(la - la1);
            if (finite && isfinite(b1.x) && isfinite(b1.y) && isfinite(b1.z)) {
                const float lb1 = 0.2126f * b1.x + 0.7152f * b1.y + 0.0722f * b1.z;
                v[4] += fabsf(lb - lb1);
            }
        }
    }
    __shared__ float red[6][256];
#pragma unroll
    for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
#pragma unroll
            for (int k = 0; k < 6; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) atomicAdd(&guard->acc[c][threadIdx.x], red[threadIdx.x][0]);
}

// Decision per camera: fall back to the filter when any output is not
// finite, when the output's mean log(1 + luminance) moved by more than the
// budget (relative) from the input's, or when the output kept more than the budgeted share of
// the input's high-frequency energy (the network did not denoise).  Clears
// the accumulators for the next frame.
__global__ void k_nn_guard_decide(NnGuard* guard, int count) {
    const int c = threadIdx.x;
    if (c >= 4) return;
    if (c < count) {
        const float* a = guard->acc[c];
        const float shift = a[1] > 0.0f ? fabsf(a[2] - a[1]) / a[1] : (a[2] != 0.0f ? 1.0f : 0.0f);
        const float noise = a[3] > 0.0f ? a[4] / a[3] : 0.0f;
        const int fb = (a[0] > 0.0f || !(shift <= guard->max_mean_shift)
                           || !(noise <= guard->max_residual_noise)) ? 1 : 0;
        guard->fallback[c] = fb;
        guard->metric[c][0] = a[0];
        guard->metric[c][1] = shift;
        guard->metric[c][2] = noise;
        if (fb) atomicAdd(&guard->frames_fallback, 1u);
    } else {
        guard->fallback[c] = 0;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) guard->acc[c][k] = 0.0f;
    if (c == 0) guard->frames_checked += 1u;
}

// linear_to_display_u8 (src/realtime/viewer/cuda_gl_presenter.cu:23-26):
// gamma 2 (square root) of the radiance clamped below at 0 (fmaxf also sends
// NaN to 0) and above at 0.999, times 255, rounded half away from zero.
// sqrtf is correctly rounded in this build, so the result is bit-exact with
// any correctly rounded restatement.
__device__ __forceinline__ uint32_t display_u8(float v) {
    const float d = fminf(sqrtf(fmaxf(v, 0.0f)), 0.999f);
    return (uint32_t)roundf(d * 255.0f);   // d * 255 lies in [0, 254.745]
}

// tone_map_to_rgba8 (cuda_gl_presenter.cu:28-38): one float4 in and one
// RGBA8 word out per pixel, alpha 255; 20 bytes of HBM traffic per pixel.
// Each thread handles four pixels 256 apart inside its block's 1024-pixel
// tile: all four loads are issued before the first store, and both sides
// are streamed with non-temporal accesses (the frame is read once and the
// output is not read back by this launch).
constexpr int kToneMapTile = 1024;

__global__ __launch_bounds__(256) void k_tone_map_rgba8(const float4* __restrict__ src,
    uint32_t* __restrict__ dst, int64_t n) {
    typedef float f4v __attribute__((ext_vector_type(4)));
    const int64_t base = (int64_t)blockIdx.x * kToneMapTile + threadIdx.x;
    f4v c[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int64_t i = base + j * 256;
        if (i < n) c[j] = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(src + i));
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int64_t i = base + j * 256;
        if (i < n) {
            const uint32_t w = display_u8(c[j].x) | (display_u8(c[j].y) << 8)
                             | (display_u8(c[j].z) << 16) | 0xFF000000u;
            __builtin_nontemporal_store(w, dst + i);
        }
    }
}

}  // namespace

namespace {

// direction_debug_kernel (programs.cu:1807-1826): the world direction of each
// pixel centre's camera ray, one byte per axis, 255 * clamp01(0.5 (v + 1))
// truncated (encode_direction_channel, programs.cu:180-182), alpha 255
__global__ void k_direction_debug(const rt_camera* cam, uint32_t* rgba, int width, int height) {
    const int x = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = (int)(blockIdx.y * blockDim.y + threadIdx.y);
    if (x >= width || y >= height) return;
    f3 o, d;
    camera_ray(*cam, (double)((float)x + 0.5f), (double)((float)y + 0.5f), &o, &d);
    auto enc = [](float v) -> uint32_t {
        const float t = 0.5f * (v + 1.0f);
        return (uint32_t)(255.0f * (t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t)));
    };
    rgba[(size_t)y * (size_t)width + (size_t)x] = enc(d.x) | (enc(d.y) << 8) | (enc(d.z) << 16) | (255u << 24);
}

}  // namespace

hipError_t wf_direction_debug(const rt_camera* cam, uint32_t* rgba, int width, int height,
    hipStream_t st) {
    const dim3 block(16, 16);
    const dim3 grid((unsigned)((width + 15) / 16), (unsigned)((height + 15) / 16));
    k_direction_debug<<<grid, block, 0, st>>>(cam, rgba, width, height);
    return hipGetLastError();
}

hipError_t wf_tone_map_rgba8(const float4* src, uint32_t* dst, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int64_t blocks = (n + kToneMapTile - 1) / kToneMapTile;
    if (blocks > INT32_MAX) return hipErrorInvalidValue;
    k_tone_map_rgba8<<<(unsigned)blocks, 256, 0, st>>>(src, dst, n);
    return hipGetLastError();
}

hipError_t wf_resolve(const CameraView& cams, const ResolveParams& rp, float4* beauty,
    float4* normal, float* depth, const float4* hbeauty, const float4* hnormal,
    const float* hdepth, float2* flow, float* trust, int total, hipStream_t st) {
    if (total <= 0) return hipSuccess;
    k_resolve<<<(total + 255) / 256, 256, 0, st>>>(cams, rp, beauty, normal, depth, hbeauty,
        hnormal, hdepth, flow, trust, total);
    return hipGetLastError();
}

hipError_t wf_nn_guard(const CameraView& cams, const float4* in, const float4* out,
    const float* depth, NnGuard* guard, hipStream_t st) {
    if (cams.count <= 0 || cams.count > 4) return hipErrorInvalidValue;
    k_nn_guard_stats<<<dim3(64, (unsigned)cams.count), 256, 0, st>>>(cams, in, out, depth, guard);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_nn_guard_decide<<<1, 64, 0, st>>>(guard, cams.count);
    return hipGetLastError();
}

hipError_t wf_denoise_filter(const CameraView& cams, int iterations, const float4* beauty,
    const float4* normal, const float4* albedo, const float* depth, float4* out,
    float4* scratch, int total, hipStream_t st, const int32_t* gate) {
    if (total <= 0) return hipSuccess;
    // passes alternate between `scratch` and `out`, the last one lands in `out`
    const float4* src = beauty;
    for (int it = 0; it < iterations; ++it) {
        float4* dst = ((iterations - 1 - it) & 1) == 0 ? out : scratch;
        k_denoise_pass<<<(total + 255) / 256, 256, 0, st>>>(cams, 1 << it, src, normal, albedo,
            depth, dst, total, gate);
        src = dst;
    }
    if (iterations <= 0) {
        if (gate != nullptr) return hipErrorInvalidValue;
        return hipMemcpyAsync(out, beauty, (size_t)total * sizeof(float4), hipMemcpyDeviceToDevice, st);
    }
    return hipGetLastError();
}

}  // namespace rtk
