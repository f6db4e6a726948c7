// rt_restir.h -- ReSTIR DI for analytic lights at the first path vertex.
//
// Follows src/realtime/gpu/programs.cu:2105-2449 and src/common/restir_di.h:
// RIS over <= 32 light candidates, temporal reuse of the reprojected pixel's
// reservoir from the previous frame, spatial reuse of previous-frame
// neighbours (the reference's rotated 8-ring, or uniform offsets in a disk of
// restir_spatial_radius pixels), BASIC bias correction that re-traces one
// ray per reuse source from the previous camera, and shading of the selected
// sample.  Random numbers: candidate k uses group 8+k (light, u0, u1, RIS
// choice); merge choices use groups 40-42, disk offsets groups 43-46.
#pragma once

#include "rt_camera.h"
#include "rt_internal.h"
#include "rt_shading.h"
#include "rt_traverse.h"

namespace rtk {

__device__ __forceinline__ bool rs_finite_pos(float v) { return v > 0.0f && isfinite(v); }
__device__ __forceinline__ float rs_clamp_unit(float v) {
    return v < 0.0f ? 0.0f : (v < 1.0f ? v : 0.99999994f);
}

__device__ __forceinline__ void rs_init(Reservoir& r) {
    r.light_index = -1;
    r.sample_u0 = 0.0f;
    r.sample_u1 = 0.0f;
    for (int a = 0; a < 3; ++a) {
        r.position[a] = 0.0f;
        r.normal[a] = 0.0f;
    }
    r.material_type = -1;
    r.primitive_type = -1;
    r.primitive_index = -1;
    r.weight_sum = 0.0f;
    r.selected_target = 0.0f;
    r.estimator_weight = 0.0f;
    r.candidate_count = 0;
    r.temporal_candidate_count = 0;
    r.spatial_candidate_count = 0;
    r.bias_correction_source_count = 0;
    r.age = 0;
    r.valid = 0;
}

// restir_update (restir_di.h:67-90)
__device__ __forceinline__ bool rs_update(Reservoir& r, int light, float u0, float u1, float target,
    float weight, int mult, float u) {
    if (light < 0 || mult <= 0) return false;
    r.candidate_count += mult;
    if (!rs_finite_pos(target) || !rs_finite_pos(weight)) return false;
    const float next = r.weight_sum + weight;
    if (!rs_finite_pos(next)) return false;
    const bool sel = rs_clamp_unit(u) * next < weight;
    if (sel) {
        r.light_index = light;
        r.sample_u0 = u0;
        r.sample_u1 = u1;
        r.selected_target = target;
        r.valid = 1;
    }
    r.weight_sum = next;
    return sel;
}

// restir_finalize (restir_di.h:92-107)
__device__ __forceinline__ void rs_finalize(Reservoir& r) {
    r.estimator_weight = 0.0f;
    if (r.valid == 0 || r.candidate_count <= 0 || !rs_finite_pos(r.selected_target)
        || !rs_finite_pos(r.weight_sum)) {
        r.valid = 0;
        return;
    }
    r.estimator_weight = r.weight_sum / ((float)r.candidate_count * r.selected_target);
    if (!rs_finite_pos(r.estimator_weight)) {
        r.valid = 0;
        r.estimator_weight = 0.0f;
    }
}

// restir_finalize_basic_bias_correction (restir_di.h:109-125)
__device__ __forceinline__ void rs_finalize_basic(Reservoir& r, float sst, float sum, int count) {
    r.estimator_weight = 0.0f;
    r.bias_correction_source_count = count > 0 ? count : 0;
    if (r.valid == 0 || count <= 0 || !rs_finite_pos(r.selected_target)
        || !rs_finite_pos(r.weight_sum) || !rs_finite_pos(sst) || !rs_finite_pos(sum)) {
        r.valid = 0;
        return;
    }
    r.estimator_weight = r.weight_sum * sst / (r.selected_target * sum);
    if (!rs_finite_pos(r.estimator_weight)) {
        r.valid = 0;
        r.estimator_weight = 0.0f;
    }
}

// restir_merge_temporal / restir_merge_spatial (restir_di.h:155-199)
__device__ __forceinline__ int rs_merge(Reservoir& cur, const Reservoir& prev, float target,
    int max_count, float u, bool temporal, bool* selected) {
    *selected = false;
    if (prev.valid == 0 || prev.candidate_count <= 0 || !rs_finite_pos(prev.estimator_weight))
        return 0;
    int mult = prev.candidate_count;
    if (max_count > 0 && mult > max_count) mult = max_count;
    const float cw = rs_finite_pos(target) ? target * prev.estimator_weight * (float)mult : 0.0f;
    *selected = rs_update(cur, prev.light_index, prev.sample_u0, prev.sample_u1, target, cw, mult, u);
    if (temporal) {
        cur.temporal_candidate_count += mult;
        cur.age = prev.age + 1;
    } else {
        cur.spatial_candidate_count += mult;
    }
    return mult;
}

// restir_temporal_surface_valid (restir_di.h:137-153)
__device__ __forceinline__ bool rs_surface_valid(const Reservoir& cur, const Reservoir& prev,
    float max_dist, float min_dot, int max_age) {
    if (prev.valid == 0 || prev.light_index < 0 || prev.age < 0 || prev.age >= max_age
        || cur.material_type != prev.material_type || cur.primitive_type != prev.primitive_type
        || cur.primitive_index != prev.primitive_index)
        return false;
    const float lim = max_dist > 0.0f ? max_dist : 0.0f;
    const float x = cur.position[0] - prev.position[0];
    const float y = cur.position[1] - prev.position[1];
    const float z = cur.position[2] - prev.position[2];
    const float d2 = x * x + y * y + z * z;
    const float nd = cur.normal[0] * prev.normal[0] + cur.normal[1] * prev.normal[1]
                     + cur.normal[2] * prev.normal[2];
    return d2 <= lim * lim && nd >= min_dot;
}

__device__ __forceinline__ void rs_surface(const HitInfo& h, Reservoir& r) {
    r.position[0] = h.pos.x;
    r.position[1] = h.pos.y;
    r.position[2] = h.pos.z;
    r.normal[0] = h.sn.x;
    r.normal[1] = h.sn.y;
    r.normal[2] = h.sn.z;
    r.material_type = h.mat_type;
    r.primitive_type = h.prim_type;
    r.primitive_index = h.prim_index;
}

// restir_unshadowed_numerator (programs.cu:2123-2145)
template <int PBR, bool PTEX>
__device__ inline f3 rs_numerator(const SceneView& s, const HitInfo& h, f3 ray_d, f3 thr,
    const LightSample& L, bool bsdf_avail, float* target) {
    *target = 0.0f;
    if (!L.valid || L.pdf <= 0.0f) return mk3(0, 0, 0);
    float bpdf = 0.0f;
    const f3 resp = direct_response<PBR, PTEX>(s, h, ray_d, L.dir, thr, &bpdf);
    if (maxc3(resp) <= 0.0f) return mk3(0, 0, 0);
    const float mis = L.delta || !bsdf_avail ? 1.0f : power_heuristic(L.pdf, bpdf);
    const f3 num = (L.emission * resp) * mis;
    *target = fmaxf(0.0f, 0.2126f * num.x + 0.7152f * num.y + 0.0722f * num.z);
    return num;
}

// restir_previous_pixel (programs.cu:2147-2171)
__device__ inline bool rs_previous_pixel(const RestirView& rv, int cam, int width, int height,
    f3 pos, int* index) {
    *index = -1;
    if (rv.history == nullptr || rv.dyn->restir_history == 0 || rv.dyn->history_length[cam] <= 0
        || rv.dyn->prev_valid[cam] == 0)
        return false;
    const rt_camera& pc = rv.prev_cams[cam];
    const f3 origin = mk3((float)pc.origin[0], (float)pc.origin[1], (float)pc.origin[2]);
    const f3 off = pos - origin;
    const f3 bx = mk3((float)pc.basis_x[0], (float)pc.basis_x[1], (float)pc.basis_x[2]);
    const f3 by = mk3((float)pc.basis_y[0], (float)pc.basis_y[1], (float)pc.basis_y[2]);
    const f3 bz = mk3((float)pc.basis_z[0], (float)pc.basis_z[1], (float)pc.basis_z[2]);
    float px, py;
    project_camera(pc, mk3(dot3(off, bx), dot3(off, by), dot3(off, bz)), &px, &py);
    if (!isfinite(px) || !isfinite(py)) return false;
    const int x = (int)floorf(px);
    const int y = (int)floorf(py);
    if (x < 0 || x >= width || y < 0 || y >= height) return false;
    *index = y * width + x;
    return true;
}

// restir_target_at_previous_surface (programs.cu:2210-2247)
template <int BLOCK, int PBR, bool PTEX>
__device__ inline float rs_target_at_source(const SceneView& s, const RestirView& rv, int cam,
    const Reservoir& src, const Reservoir& sel, bool bsdf_avail, const TraceStack& st) {
    if (rv.dyn->prev_valid[cam] == 0) return 0.0f;
    const rt_camera& pc = rv.prev_cams[cam];
    const f3 sp = mk3(src.position[0], src.position[1], src.position[2]);
    const f3 origin = mk3((float)pc.origin[0], (float)pc.origin[1], (float)pc.origin[2]);
    const f3 off = sp - origin;
    const float dist = len3(off);
    if (!isfinite(dist) || dist <= kRayEps) return 0.0f;
    const f3 dir = off * (1.0f / dist);
    const TraceResult tr = trace_wide<BLOCK, false>(s, origin, dir, kRayEps, dist + 0.05f, st);
    if (!tr.found) return 0.0f;
    HitInfo h;
    make_hit(s, origin, dir, tr.t, (int)tr.prim, h);
    if (h.mat_type != src.material_type || h.prim_type != src.primitive_type
        || h.prim_index != src.primitive_index)
        return 0.0f;
    const float tol = fmaxf(0.01f, 0.02f * dist);
    if (lensq3(h.pos - sp) > tol * tol
        || dot3(h.sn, mk3(src.normal[0], src.normal[1], src.normal[2])) < 0.95f)
        return 0.0f;
    const LightSample L = analytic_candidate(s, h.pos, sel.light_index, sel.sample_u0, sel.sample_u1);
    float target = 0.0f;
    // a fresh path state at the previous surface: unit throughput
    rs_numerator<PBR, PTEX>(s, h, dir, mk3(1.0f, 1.0f, 1.0f), L, bsdf_avail, &target);
    return target;
}

struct RestirOut {
    bool shadow;
    f3 origin, dir, numerator;
    float max_t;
    float W;
};

// accumulate_restir_analytic_direct_light (programs.cu:2315-2449) up to the
// shadow ray, which the connect stage traces.  `cam` is the camera's slot and
// `base` its first pixel in the reservoir buffers, which hold all cameras
// back to back; `key` is the camera key of the slot's random numbers and
// `thr` the path throughput at the vertex.  PTEX: the scene binds OpenPBR
// inputs to textures.
template <int BLOCK, int PBR, bool PTEX>
__device__ inline RestirOut restir_shade(const SceneView& s, const RestirView& rv,
    const HitInfo& h, f3 ray_d, f3 thr, int cam, uint32_t key, int base, int x, int y, int width, int height,
    uint32_t pixel, uint32_t sample, uint32_t stream, bool bsdf_avail, Reservoir* out,
    const TraceStack& st) {
    RestirOut ro{};
    ro.shadow = false;
    if (h.mat_type == RT_MAT_DIFFUSE_LIGHT || s.analytic_count <= 0) return ro;
    Reservoir r;
    rs_init(r);
    const int req = rv.initial_candidates > 0 ? rv.initial_candidates : 1;
    const int ncand = req < 32 ? req : 32;
    for (int k = 0; k < ncand; ++k) {
        const float4 q = rng4(pixel, sample, key, 0u, kGroupRestir0 + (uint32_t)k, stream);
        const int li = pick_analytic(s, q.x);
        const LightSample L = analytic_candidate(s, h.pos, li, q.y, q.z);
        float target = 0.0f;
        rs_numerator<PBR, PTEX>(s, h, ray_d, thr, L, bsdf_avail, &target);
        const float w = L.pdf > 0.0f ? target / L.pdf : 0.0f;
        rs_update(r, li, q.y, q.z, target, w, 1, q.w);
    }
    const int canonical = r.candidate_count;
    float sel_u[12], off_u[16];
    for (int g = 0; g < 3; ++g) {
        const float4 q = rng4(pixel, sample, key, 0u, kGroupRestir0 + 32u + (uint32_t)g, stream);
        sel_u[4 * g] = q.x; sel_u[4 * g + 1] = q.y; sel_u[4 * g + 2] = q.z; sel_u[4 * g + 3] = q.w;
    }
    for (int g = 0; g < 4; ++g) {
        const float4 q = rng4(pixel, sample, key, 0u, kGroupRestir0 + 35u + (uint32_t)g, stream);
        off_u[4 * g] = q.x; off_u[4 * g + 1] = q.y; off_u[4 * g + 2] = q.z; off_u[4 * g + 3] = q.w;
    }
    Reservoir cur;
    rs_init(cur);
    rs_surface(h, cur);
    int src_idx[9];
    int mults[9];
    int nsrc = 0, selected_source = -1;
    if (rv.temporal_reuse != 0) {
        int pidx;
        if (rs_previous_pixel(rv, cam, width, height, h.pos, &pidx)) {
            pidx += base;
            const Reservoir prev = rv.history[pidx];
            if (rs_surface_valid(cur, prev, fmaxf(0.005f, 0.01f * h.t), 0.95f, rv.max_history_age)) {
                const LightSample L = analytic_candidate(s, h.pos, prev.light_index, prev.sample_u0, prev.sample_u1);
                float target = 0.0f;
                rs_numerator<PBR, PTEX>(s, h, ray_d, thr, L, bsdf_avail, &target);
                bool sel = false;
                const int mult = rs_merge(r, prev, target, rv.max_temporal_candidates, sel_u[0], true, &sel);
                if (mult > 0) {
                    src_idx[nsrc] = pidx;
                    mults[nsrc] = mult;
                    if (sel) selected_source = nsrc;
                    ++nsrc;
                }
            }
        }
    }
    const int nsp = rv.spatial_neighbors < 8 ? rv.spatial_neighbors : 8;
    if (nsp > 0) {
        int ridx;
        if (rs_previous_pixel(rv, cam, width, height, h.pos, &ridx)) {
            const int rot = (x * 3 + y * 5 + (int)stream) & 7;
            const int bx = ridx % width, by = ridx / width;
            for (int j = 0; j < nsp; ++j) {
                int dx, dy;
                if (rv.spatial_radius > 1) {
                    const float rad = (float)rv.spatial_radius * sqrtf(off_u[2 * j]);
                    float sp_, cp_;
                    sincos_turn(off_u[2 * j + 1], &sp_, &cp_);
                    dx = (int)floorf(rad * cp_ + 0.5f);
                    dy = (int)floorf(rad * sp_ + 0.5f);
                } else {
                    const int k = (rot + j) & 7;
                    dx = k == 0 ? -1 : (k == 1 ? 1 : (k == 2 || k == 3 ? 0 : (k == 4 || k == 6 ? -1 : 1)));
                    dy = k < 2 ? 0 : (k == 2 ? -1 : (k == 3 ? 1 : (k < 6 ? -1 : 1)));
                }
                const int nx = bx + dx, ny = by + dy;
                if (nx < 0 || nx >= width || ny < 0 || ny >= height) continue;
                const int nidx = base + ny * width + nx;
                const Reservoir prev = rv.history[nidx];
                if (!rs_surface_valid(cur, prev, fmaxf(0.01f, 0.02f * h.t), 0.95f, rv.max_history_age)) continue;
                const LightSample L = analytic_candidate(s, h.pos, prev.light_index, prev.sample_u0, prev.sample_u1);
                float target = 0.0f;
                rs_numerator<PBR, PTEX>(s, h, ray_d, thr, L, bsdf_avail, &target);
                bool sel = false;
                const int mult = rs_merge(r, prev, target, rv.max_spatial_candidates, sel_u[1 + j], false, &sel);
                if (mult > 0 && nsrc < 9) {
                    src_idx[nsrc] = nidx;
                    mults[nsrc] = mult;
                    if (sel) selected_source = nsrc;
                    ++nsrc;
                }
            }
        }
    }
    rs_surface(h, r);
    if (rv.bias_correction == 1 && nsrc > 0 && r.valid != 0) {
        float sst = r.selected_target;
        float sum = r.selected_target * (float)canonical;
        for (int i = 0; i < nsrc; ++i) {
            const float t = rs_target_at_source<BLOCK, PBR, PTEX>(s, rv, cam, rv.history[src_idx[i]], r,
                bsdf_avail, st);
            sum += t * (float)mults[i];
            if (selected_source == i) sst = t;
        }
        rs_finalize_basic(r, sst, sum, nsrc);
    } else {
        rs_finalize(r);
    }
    if (out != nullptr) *out = r;
    if (r.valid == 0) return ro;
    const LightSample L = analytic_candidate(s, h.pos, r.light_index, r.sample_u0, r.sample_u1);
    float target = 0.0f;
    const f3 num = rs_numerator<PBR, PTEX>(s, h, ray_d, thr, L, bsdf_avail, &target);
    if (!L.valid || target <= 0.0f) return ro;
    const float max_t = L.infinite ? kRayFar : L.dist - kRayEps * 3.0f;
    if (max_t <= kRayEps) return ro;
    ro.shadow = true;
    ro.origin = h.pos + L.dir * (kRayEps * 2.0f);
    ro.dir = L.dir;
    ro.max_t = max_t;
    ro.numerator = num;
    ro.W = r.estimator_weight;
    return ro;
}

}  // namespace rtk
