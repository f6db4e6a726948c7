// rt_shading.h -- device surface interaction, light sampling and BSDF
// sampling for the shade stage of the wavefront pipeline.
//
// Semantics follow the reference megakernel (src/realtime/gpu/programs.cu):
//   populate_material :961-1009          emission MIS   :571-647
//   sample_direct_light :1057-1232       direct response :1026-1055
//   accumulate_direct_light :2249-2280   sample_bsdf    :2451-2565
// with two structural changes: the light owning a primitive is found through
// a prim -> light table instead of the O(L) scan of find_packed_light
// (:559-568), and light selection uses a binary search of the CDF (the same
// index as the linear walk of light_sampling.h:90-102, since the CDF is
// non-decreasing).  Random numbers come from the counter RNG, two 4-tuples
// per bounce; rejection loops are replaced by closed-form samplers.
#pragma once

#include "rt_internal.h"
#include "rt_math.h"
#include "rt_openpbr.h"
#include "rt_texture.h"

namespace rtk {

struct HitInfo {
    float t;
    f3 pos, gn, sn, base, emission;
    float fuzz, ior;
    float tu, tv;   // texture coordinates (set where a texture is bound)
    int mat_type, openpbr, prim, prim_type, prim_index;
    bool front;
};

__device__ __forceinline__ f3 background_at(const SceneView& s, f3 d) {
    if (s.background_mode == RT_BG_SKY_GRADIENT) {
        const f3 dn = normalize3(d);
        const float t = 0.5f * (dn.y + 1.0f);
        return mk3(s.bg[0], s.bg[1], s.bg[2]) * (1.0f - t) + mk3(s.bg_top[0], s.bg_top[1], s.bg_top[2]) * t;
    }
    return mk3(s.bg[0], s.bg[1], s.bg[2]);
}

__device__ __forceinline__ void set_face(f3 d, f3 outward, HitInfo& h) {
    h.front = dot3(d, outward) < 0.0f;
    h.gn = outward;
    h.sn = h.front ? outward : -outward;
}

// A hit record's primitive id: >= 0 a surface primitive, -1 none, and
// kMediumHit - i (<= -2) a scatter event inside medium i.
constexpr int kMediumHit = -2;

// TEX = false leaves textures out (a scene without them): albedo and
// emission textures and the OpenPBR base and emission colour bindings
template <bool TEX = true>
__device__ inline void make_hit(const SceneView& s, f3 o, f3 d, float t, int prim, HitInfo& h) {
    h.t = t;
    h.tu = 0.0f;
    h.tv = 0.0f;
    int mat;
    if (prim <= kMediumHit) {
        // the medium candidate of intersect_scene (programs.cu:1736-1748): no
        // primitive, front face, normals against the ray, texture coordinates 0
        const rt_medium md = s.media[kMediumHit - prim];
        h.prim = -1;
        h.prim_type = -1;
        h.prim_index = -1;
        h.pos = o + d * t;
        h.front = true;
        h.gn = normalize3(d) * -1.0f;
        h.sn = h.gn;
        mat = md.material_index;
    } else if (prim < s.sphere_count) {
        h.prim = prim;
        const rt_sphere sp = s.spheres[prim];
        h.prim_type = 0;
        h.prim_index = prim;
        h.pos = o + d * t;
        set_face(d, div3(h.pos - ld3(sp.center), sp.radius), h);
        mat = sp.material;
    } else if (prim < s.sphere_count + s.quad_count) {
        h.prim = prim;
        const int qi = prim - s.sphere_count;
        const rt_quad q = s.quads[qi];
        h.prim_type = 1;
        h.prim_index = qi;
        const f3 n = cross3(ld3(q.edge_u), ld3(q.edge_v));
        h.pos = o + d * t;
        set_face(d, div3(n, sqrtf(lensq3(n))), h);
        mat = q.material;
    } else {
        h.prim = prim;
        const int ti = prim - s.sphere_count - s.quad_count;
        const rt_triangle tr = s.triangles[ti];
        h.prim_type = 2;
        h.prim_index = ti;
        const f3 p0 = ld3(tr.p0);
        h.pos = o + d * t;
        const f3 normal = normalize3(cross3(ld3(tr.p1) - p0, ld3(tr.p2) - p0));
        set_face(d, normal, h);
        mat = tr.material;
        if (s.tri_n != nullptr && s.tri_n[9 * (int64_t)ti] == s.tri_n[9 * (int64_t)ti]) {
            // interpolated vertex normals (programs.cu:1535-1545)
            const float* vn = s.tri_n + 9 * (int64_t)ti;
            const f3 e1 = ld3(tr.p1) - p0, e2 = ld3(tr.p2) - p0;
            const f3 pvec = cross3(d, e2);
            const float inv_det = 1.0f / dot3(e1, pvec);
            const f3 tvec = o - p0;
            const float bu = dot3(tvec, pvec) * inv_det;
            const float bv = dot3(d, cross3(tvec, e1)) * inv_det;
            const float bw = 1.0f - bu - bv;
            f3 sh = ld3(vn) * bw + (ld3(vn + 3) * bu + ld3(vn + 6) * bv);
            if (!(fabsf(sh.x) < 1e-7f && fabsf(sh.y) < 1e-7f && fabsf(sh.z) < 1e-7f)) {
                sh = normalize3(sh);
                if (dot3(sh, normal) < 0.0f) sh = sh * -1.0f;
                h.sn = h.front ? sh : sh * -1.0f;
            }
        }
    }
    if (mat < 0 || mat >= s.material_count) {
        h.base = mk3(0, 0, 0); h.emission = mk3(0, 0, 0); h.mat_type = 0; h.ior = 1.0f;
        h.fuzz = 0.0f; h.openpbr = -1;
        return;
    }
    const rt_material m = s.materials[mat];
    h.openpbr = m.openpbr_index;
    h.mat_type = m.type;
    h.fuzz = m.fuzz;
    h.ior = m.ior;
    h.base = ld3(m.albedo);
    h.emission = ld3(m.emission);
    const bool pbr_ok = m.type == RT_MAT_OPENPBR && m.openpbr_index >= 0 && m.openpbr_index < s.openpbr_count;
    if (TEX) {
        // populate_material's albedo and emission textures (programs.cu:978-981;
        // an emission texture only shows on the front of a diffuse light) and
        // the OpenPBR bindings read at every hit (:995-1008)
        const int at = s.mat_albedo_tex != nullptr ? s.mat_albedo_tex[mat] : -1;
        const int et = s.mat_emission_tex != nullptr && m.type == RT_MAT_DIFFUSE_LIGHT && h.front
                           ? s.mat_emission_tex[mat] : -1;
        const bool pt = s.pbr_tex != nullptr && pbr_ok;
        if ((at >= 0 || et >= 0 || pt) && h.prim_type >= 0)
            surface_uv(s, o, d, h.prim_type, h.prim_index, h.pos, &h.tu, &h.tv);
        if (at >= 0) h.base = eval_texture(s, at, h.tu, h.tv, h.pos);
        if (et >= 0) h.emission = eval_texture(s, et, h.tu, h.tv, h.pos);
    }
    if (h.mat_type == RT_MAT_DIELECTRIC) { h.base = mk3(0.92f, 0.95f, 1.0f); h.emission = mk3(0, 0, 0); }
    if (h.mat_type == RT_MAT_DIFFUSE_LIGHT && !h.front) h.emission = mk3(0, 0, 0);
    if (h.mat_type != RT_MAT_DIFFUSE_LIGHT) h.emission = mk3(0, 0, 0);
    if (h.mat_type == RT_MAT_OPENPBR) {
        if (!pbr_ok) {
            h.mat_type = -1; h.base = mk3(0, 0, 0); h.emission = mk3(0, 0, 0);
            return;
        }
        const rt_openpbr_material& p = s.openpbr[m.openpbr_index];
        if (TEX && s.pbr_tex != nullptr) {
            const rt_openpbr_textures tb = s.pbr_tex[m.openpbr_index];
            opbr::MatTex<opbr::MatMem> mt(opbr::MatMem{&p});
            if (tb.base_color >= 0)
                mt.bc = source_to_linear(eval_texture(s, tb.base_color, h.tu, h.tv, h.pos), tb.base_color_space);
            if (tb.emission_color >= 0)
                mt.ec = source_to_linear(eval_texture(s, tb.emission_color, h.tu, h.tv, h.pos),
                    tb.emission_color_space);
            h.base = mt.bc;
            h.emission = openpbr_emission_v(mt);
        } else {
            h.base = ld3(p.base_color);
            h.emission = openpbr_emission(p);
        }
    }
}

// evaluate_openpbr_scattering_material (programs.cu:944-959): the record with
// its texture-bound inputs replaced by the textures' values at the hit (the
// base colour is the hit's, already textured)
template <class M>
__device__ inline opbr::MatTex<M> pbr_scattering(const SceneView& s, const M& m, const HitInfo& h) {
    opbr::MatTex<M> mt(m);
    const rt_openpbr_textures tb = s.pbr_tex[h.openpbr];
    mt.bc = h.base;
    if (tb.specular_color >= 0)
        mt.sc = source_to_linear(eval_texture(s, tb.specular_color, h.tu, h.tv, h.pos), tb.specular_color_space);
    if (tb.transmission_color >= 0)
        mt.tc = source_to_linear(eval_texture(s, tb.transmission_color, h.tu, h.tv, h.pos),
            tb.transmission_color_space);
    if (tb.base_metalness >= 0) mt.metal = eval_texture(s, tb.base_metalness, h.tu, h.tv, h.pos).x;
    if (tb.specular_roughness >= 0) mt.rough = eval_texture(s, tb.specular_roughness, h.tu, h.tv, h.pos).x;
    return mt;
}

// path-vertex state carried between bounces
struct PathPrev {
    f3 pos, n;
    float pdf;
    int prim;      // -1: none
    int mat;       // material type
    bool valid, delta;
};

__device__ __forceinline__ float power_heuristic(float a, float b) {
    if (a <= 0.0f) return 0.0f;
    const float a2 = a * a, b2 = b * b;
    return a2 / (a2 + b2);
}
__device__ __forceinline__ float cone_pdf(float cos_max) {
    const float sa = 6.28318530717958648f * (1.0f - cos_max);
    return sa > 1e-12f ? 1.0f / sa : 0.0f;
}
__device__ __forceinline__ float area_pdf(float area, float dsq, float acos_) {
    if (area <= 1e-12f || dsq <= 0.0f || acos_ <= 1e-8f) return 0.0f;
    return dsq / (area * acos_);
}

__device__ __forceinline__ float eff_selection_pdf(const SceneView& s, int light, int excluded_prim) {
    const int ex = excluded_prim < 0 ? -1 : s.light_of_prim[excluded_prim];
    if (ex == light) return 0.0f;
    const float elig = ex < 0 ? 1.0f : 1.0f - s.lights[ex].selection_pdf;
    return elig > 1e-8f ? s.lights[light].selection_pdf / elig : 0.0f;
}

__device__ inline float emission_mis(const SceneView& s, const PathPrev& pv, f3 ray_d,
    const HitInfo* h, bool env) {
    if (!pv.valid || pv.delta) return 1.0f;
    if (!env && h->prim < 0) return 1.0f;   // a medium event: no primitive, no light
    const int light = env ? s.env_light : s.light_of_prim[h->prim];
    if (light < 0) return 1.0f;
    const f3 dir = normalize3(ray_d);
    float cond;
    if (env) {
        cond = pv.mat == RT_MAT_ISOTROPIC ? kUniformSpherePdf : fmaxf(dot3(pv.n, dir), 0.0f) / kPi;
    } else if (!h->front) {
        cond = 0.0f;
    } else if (h->prim_type == 0) {
        const rt_sphere sp = s.spheres[h->prim_index];
        const float dsq = lensq3(ld3(sp.center) - pv.pos);
        const float rsq = sp.radius * sp.radius;
        cond = dsq <= rsq ? 0.0f : cone_pdf(sqrtf(fmaxf(0.0f, 1.0f - rsq / dsq)));
    } else {
        const float dsq = lensq3(h->pos - pv.pos);
        const float ac = fmaxf(dot3(h->gn, dir * -1.0f), 0.0f);
        float area;
        if (h->prim_type == 1) {
            const rt_quad q = s.quads[h->prim_index];
            area = len3(cross3(ld3(q.edge_u), ld3(q.edge_v)));
        } else {
            const rt_triangle t = s.triangles[h->prim_index];
            area = 0.5f * len3(cross3(ld3(t.p1) - ld3(t.p0), ld3(t.p2) - ld3(t.p0)));
        }
        cond = area_pdf(area, dsq, ac);
    }
    const float sel = eff_selection_pdf(s, light, pv.prim);
    return power_heuristic(pv.pdf, sel * cond);
}

struct LightSample {
    f3 dir, emission;
    float dist, pdf;
    bool infinite, valid, delta;
};

__device__ __forceinline__ f3 cone_dir(f3 axis, float cos_max, float u0, float u1) {
    const float ct = 1.0f - u0 * (1.0f - cos_max);
    const float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    float sp, cp;
    sincos_turn(u1, &sp, &cp);
    const f3 helper = fabsf(axis.x) > 0.9f ? mk3(0.0f, 1.0f, 0.0f) : mk3(1.0f, 0.0f, 0.0f);
    const f3 tangent = normalize3(cross3(helper, axis));
    const f3 bitangent = cross3(axis, tangent);
    return normalize3(axis * ct + (tangent * (st * cp) + bitangent * (st * sp)));
}

// first index with u < cdf[i] (the reference's linear walk), by bisection
__device__ __forceinline__ int pick_light(const SceneView& s, float u) {
    int lo = 0, hi = s.light_count - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (u < s.lights[mid].cdf) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// TEX = false: every emission is the material's constant (a scene without
// emission textures or OpenPBR emission bindings)
template <bool TEX = true>
__device__ inline LightSample sample_light(const SceneView& s, const HitInfo& h, float4 r) {
    LightSample ls{};
    ls.valid = false;
    const int ex = h.prim >= 0 ? s.light_of_prim[h.prim] : -1;
    const float elig = ex < 0 ? 1.0f : 1.0f - s.lights[ex].selection_pdf;
    if (elig <= 1e-8f) return ls;
    int li = -1;
    if (ex < 0) {
        const float u = r.x < 0.0f ? 0.0f : (r.x < 1.0f ? r.x : 0.99999994f);
        li = pick_light(s, u);
    } else {
        const float target = r.x * elig;
        float acc = 0.0f;
        for (int i = 0; i < s.light_count; ++i) {
            if (i == ex) continue;
            acc += s.lights[i].selection_pdf;
            if (target < acc) { li = i; break; }
        }
        if (li < 0)
            for (int i = s.light_count - 1; i >= 0; --i)
                if (i != ex) { li = i; break; }
    }
    if (li < 0) return ls;
    const LightRec L = s.lights[li];
    const float sel = L.selection_pdf / elig;
    if (sel <= 0.0f) return ls;
    const f3 sp = h.pos;
    if (L.type == 3) {
        if (h.mat_type == RT_MAT_ISOTROPIC) {
            ls.dir = sphere_dir(r.y, r.z);
            ls.pdf = sel * kUniformSpherePdf;
        } else {
            ls.dir = h.sn + sphere_dir(r.y, r.z);
            if (near_zero3(ls.dir)) ls.dir = h.sn;
            ls.dir = normalize3(ls.dir);
            const float cs = fmaxf(dot3(h.sn, ls.dir), 0.0f);
            ls.pdf = sel * cs / kPi;
        }
        ls.emission = background_at(s, ls.dir);
        ls.dist = kRayFar;
        ls.infinite = true;
        ls.valid = ls.pdf > 0.0f;
        return ls;
    }
    ls.infinite = false;
    float cond = 0.0f;
    int mat = -1;
    // the sampled point and its texture coordinates (programs.cu:1154,1166-1167,1193-1199)
    f3 lpos = mk3(0, 0, 0);
    float tu = 0.0f, tv = 0.0f;
    if (L.type == 0) {
        const rt_sphere sph = s.spheres[L.prim];
        const f3 center = ld3(sph.center);
        const f3 tc = center - sp;
        const float dsq = lensq3(tc);
        const float rsq = sph.radius * sph.radius;
        if (dsq <= rsq * (1.0f + 1e-6f)) return ls;
        const float cmax = sqrtf(fmaxf(0.0f, 1.0f - rsq / dsq));
        ls.dir = cone_dir(normalize3(tc), cmax, r.y, r.z);
        const f3 oc = sp - center;
        const float hb = dot3(oc, ls.dir);
        const float disc = hb * hb - (lensq3(oc) - rsq);
        if (disc <= 0.0f) return ls;
        ls.dist = -hb - sqrtf(disc);
        if (ls.dist <= kRayEps) return ls;
        mat = sph.material;
        cond = cone_pdf(cmax);
        if (TEX) {
            lpos = sp + ls.dir * ls.dist;
            sphere_uv(normalize3(lpos - center), &tu, &tv);
        }
    } else if (L.type == 1) {
        const rt_quad q = s.quads[L.prim];
        const f3 eu = ld3(q.edge_u), ev = ld3(q.edge_v);
        const f3 pos = ld3(q.origin) + (eu * r.y + ev * r.z);
        const f3 an = cross3(eu, ev);
        const float area = len3(an);
        if (area <= 1e-12f) return ls;
        const f3 normal = div3(an, area);
        mat = q.material;
        lpos = pos;
        tu = r.y;
        tv = r.z;
        const f3 tl = pos - sp;
        const float dsq = lensq3(tl);
        if (dsq <= 1e-12f) return ls;
        ls.dist = sqrtf(dsq);
        ls.dir = div3(tl, ls.dist);
        cond = area_pdf(area, dsq, dot3(normal, ls.dir * -1.0f));
    } else {
        const rt_triangle t = s.triangles[L.prim];
        const f3 p0 = ld3(t.p0), p1 = ld3(t.p1), p2 = ld3(t.p2);
        const float su = sqrtf(r.y < 0.0f ? 0.0f : (r.y < 1.0f ? r.y : 1.0f));
        const float w0 = 1.0f - su, w1 = su * (1.0f - r.z), w2 = su * r.z;
        const f3 pos = p0 * w0 + (p1 * w1 + p2 * w2);
        const f3 an = cross3(p1 - p0, p2 - p0);
        const float da = len3(an);
        if (da <= 1e-12f) return ls;
        const f3 normal = div3(an, da);
        mat = t.material;
        lpos = pos;
        tu = w1;
        tv = w2;
        if (TEX && s.tri_uv != nullptr) {
            const float* tc = s.tri_uv + 6 * (int64_t)L.prim;
            if (tc[0] == tc[0]) {   // not NaN: has_texcoords
                tu = tc[0] * w0 + tc[2] * w1 + tc[4] * w2;
                tv = tc[1] * w0 + tc[3] * w1 + tc[5] * w2;
            }
        }
        const f3 tl = pos - sp;
        const float dsq = lensq3(tl);
        if (dsq <= 1e-12f) return ls;
        ls.dist = sqrtf(dsq);
        ls.dir = div3(tl, ls.dist);
        cond = area_pdf(0.5f * da, dsq, dot3(normal, ls.dir * -1.0f));
    }
    if (cond <= 0.0f || mat < 0 || mat >= s.material_count) return ls;
    const rt_material m = s.materials[mat];
    // evaluate_material_emission (programs.cu:1013-1024) at the sampled point
    int et = -1;
    if (TEX) {
        if (m.type == RT_MAT_OPENPBR && s.pbr_tex != nullptr && m.openpbr_index >= 0
            && m.openpbr_index < s.openpbr_count)
            et = s.pbr_tex[m.openpbr_index].emission_color;
        else if (m.type != RT_MAT_OPENPBR && s.mat_emission_tex != nullptr)
            et = s.mat_emission_tex[mat];
    }
    if (m.type == RT_MAT_OPENPBR) {
        if (m.openpbr_index < 0 || m.openpbr_index >= s.openpbr_count) {
            ls.emission = mk3(0, 0, 0);
        } else if (TEX && et >= 0) {
            const rt_openpbr_material& p = s.openpbr[m.openpbr_index];
            opbr::MatTex<opbr::MatMem> mt(opbr::MatMem{&p});
            mt.ec = source_to_linear(eval_texture(s, et, tu, tv, lpos), s.pbr_tex[m.openpbr_index].emission_color_space);
            ls.emission = openpbr_emission_v(mt);
        } else {
            ls.emission = openpbr_emission(s.openpbr[m.openpbr_index]);
        }
    } else if (TEX && et >= 0) {
        ls.emission = eval_texture(s, et, tu, tv, lpos);
    } else {
        ls.emission = ld3(m.emission);
    }
    ls.pdf = sel * cond;
    ls.valid = ls.pdf > 0.0f && maxc3(ls.emission) > 0.0f;
    return ls;
}

// PBR selects the OpenPBR code: kPbrOff leaves it out (a response of 0: the
// shade stage's legacy-material pass, which hands OpenPBR hits to a second
// pass), kPbrCore the opaque core (scenes whose materials use no other
// lobe), kPbrFull every lobe.  thr: the path throughput, the transport
// context of the OpenPBR lobes.
enum : int { kPbrOff = 0, kPbrCore = 1, kPbrFull = 2 };

// an OpenPBR material as the scattering code reads it: the record itself, or
// with PTEX (a scene that binds OpenPBR inputs to textures) the record with
// the bound inputs evaluated at the hit
template <bool PTEX, class M, class F>
__device__ __forceinline__ void with_pbr_material(const SceneView& s, const M& m, const HitInfo& h, F&& f) {
    if constexpr (PTEX) {
        if (s.pbr_tex != nullptr) {
            f(pbr_scattering(s, m, h));
            return;
        }
    }
    f(m);
}

template <int PBR = kPbrFull, bool PTEX = false>
__device__ inline f3 direct_response(const SceneView& s, const HitInfo& h, f3 ray_d, f3 ldir,
    f3 thr, float* pdf) {
    *pdf = 0.0f;
    if (h.mat_type == RT_MAT_LAMBERTIAN) {
        const float cs = fmaxf(dot3(h.sn, ldir), 0.0f);
        *pdf = cs / kPi;
        return h.base * *pdf;
    }
    if (h.mat_type == RT_MAT_ISOTROPIC) {
        *pdf = kUniformSpherePdf;
        return h.base * *pdf;
    }
    if constexpr (PBR == kPbrOff) {
        return mk3(0, 0, 0);
    } else {
        if (h.mat_type != RT_MAT_OPENPBR) return mk3(0, 0, 0);
        f3 value;
        const rt_openpbr_material* pm = s.openpbr + h.openpbr;
        auto eval = [&](const auto& m) { openpbr_eval_v(m, h.gn, -normalize3(ray_d), ldir, thr, &value, pdf); };
        if constexpr (PBR == kPbrCore) with_pbr_material<PTEX>(s, opbr::MatCore(pm), h, eval);
        else with_pbr_material<PTEX>(s, opbr::MatMem{pm}, h, eval);
        const float cs = fabsf(dot3(h.sn, ldir));
        return value * cs;
    }
}

// ---- analytic lights (PackedAnalyticLight, light_sampling.h:30-62)
__device__ __forceinline__ bool alight_finite(int type) {
    return type == RT_ALIGHT_SPHERE || type == RT_ALIGHT_DISK || type == RT_ALIGHT_RECT
           || type == RT_ALIGHT_CYLINDER;
}

// analytic_light_normal (programs.cu:655-659): -(basis_x x basis_y)
__device__ __forceinline__ f3 alight_normal(const rt_analytic_light& L) {
    return normalize3(cross3(ld3(L.basis_x), ld3(L.basis_y)) * -1.0f);
}

// analytic_plane_coordinates (programs.cu:661-677): the point's coordinates
// in the (basis_x, basis_y) frame by the normal equations
__device__ __forceinline__ bool alight_plane_xy(const rt_analytic_light& L, f3 p, float* lx, float* ly) {
    const f3 bx = ld3(L.basis_x), by = ld3(L.basis_y);
    const f3 local = p - ld3(L.position);
    const float xx = dot3(bx, bx), xy = dot3(bx, by), yy = dot3(by, by);
    const float det = xx * yy - xy * xy;
    if (det <= 1e-16f) return false;
    const float px = dot3(local, bx), py = dot3(local, by);
    *lx = (px * yy - py * xy) / det;
    *ly = (py * xx - px * xy) / det;
    return true;
}

// try_intersect_analytic_light (programs.cu:679-793): front faces of the
// sphere, disk / rect and the cylinder's side; the nearer root in
// (tmin, tmax), else the farther
__device__ inline bool alight_hit(const rt_analytic_light& L, f3 o, f3 d, float tmin, float tmax,
    float* t, f3* pos, f3* n) {
    const f3 center = ld3(L.position);
    if (L.type == RT_ALIGHT_SPHERE) {
        const float radius = L.radius * len3(ld3(L.basis_x));
        const f3 oc = o - center;
        const float a = lensq3(d);
        const float hb = dot3(oc, d);
        const float c = lensq3(oc) - radius * radius;
        const float disc = hb * hb - a * c;
        if (a <= 1e-16f || disc < 0.0f) return false;
        const float rt = sqrtf(disc);
        float root = (-hb - rt) / a;
        if (root <= tmin || root >= tmax) root = (-hb + rt) / a;
        if (root <= tmin || root >= tmax) return false;
        const f3 p = o + d * root;
        const f3 nn = normalize3(p - center);
        if (dot3(d, nn) >= 0.0f) return false;
        *t = root;
        *pos = p;
        *n = nn;
        return true;
    }
    if (L.type == RT_ALIGHT_DISK || L.type == RT_ALIGHT_RECT) {
        const f3 nn = alight_normal(L);
        const float den = dot3(d, nn);
        if (den >= -1e-8f) return false;
        const float tt = dot3(center - o, nn) / den;
        if (tt <= tmin || tt >= tmax) return false;
        const f3 p = o + d * tt;
        float lx = 0.0f, ly = 0.0f;
        if (!alight_plane_xy(L, p, &lx, &ly)) return false;
        const bool inside = L.type == RT_ALIGHT_DISK
                                ? lx * lx + ly * ly <= L.radius * L.radius
                                : fabsf(lx) <= 0.5f * L.width && fabsf(ly) <= 0.5f * L.height;
        if (!inside) return false;
        *t = tt;
        *pos = p;
        *n = nn;
        return true;
    }
    if (L.type != RT_ALIGHT_CYLINDER) return false;
    const f3 bx = ld3(L.basis_x), by = ld3(L.basis_y), bz = ld3(L.basis_z);
    const float scale = len3(bx);
    if (scale <= 1e-8f) return false;
    const f3 ax = div3(bx, scale), ay = div3(by, scale), az = normalize3(bz);
    const f3 lo = o - center;
    const float ox = dot3(lo, ax), oy = dot3(lo, ay), oz = dot3(lo, az);
    const float dx = dot3(d, ax), dy = dot3(d, ay), dz = dot3(d, az);
    const float radius = L.radius * scale;
    const float a = dx * dx + dy * dy;
    const float hb = ox * dx + oy * dy;
    const float c = ox * ox + oy * oy - radius * radius;
    const float disc = hb * hb - a * c;
    if (a <= 1e-16f || disc < 0.0f) return false;
    const float rt = sqrtf(disc);
    float root = (-hb - rt) / a;
    float z = oz + root * dz;
    const float half_len = 0.5f * L.length * len3(bz);
    if (root <= tmin || root >= tmax || fabsf(z) > half_len) {
        root = (-hb + rt) / a;
        z = oz + root * dz;
    }
    if (root <= tmin || root >= tmax || fabsf(z) > half_len) return false;
    const float x = ox + root * dx, y = oy + root * dy;
    const f3 nn = normalize3(ax * x + ay * y);
    if (dot3(d, nn) >= 0.0f) return false;
    *t = root;
    *pos = o + d * root;
    *n = nn;
    return true;
}

// intersect_analytic_lights (programs.cu:795-813): the nearest front-facing
// hit of the finite lights that are not points, lowest index on equal
// distance.  The reference walks every light with a shrinking range; a
// light's hit distance does not depend on the range's end as long as it
// lies inside, so the walk's answer is the smallest (t, index) over the
// lights hit in (tmin, tmax), which a walk of the light BVH finds in any
// order (each light tested against the full range, boxes culled only when
// they start beyond the best hit).
__device__ __forceinline__ float light_inv(float x) {
    return 1.0f / (fabsf(x) > 1e-20f ? x : copysignf(1e-20f, x));
}

__device__ inline int intersect_analytic(const SceneView& s, f3 o, f3 d, float tmin, float tmax) {
    if (s.a_node_count <= 0) return -1;
    const f3 inv = mk3(light_inv(d.x), light_inv(d.y), light_inv(d.z));
    float best = tmax;
    int found = -1;
    // stackless depth-first walk: descend into a box the ray meets before
    // the best hit, otherwise (and after a leaf) jump past the subtree
    int k = 0;
    while (k < s.a_node_count) {
        const float4 lo = s.a_nodes[2 * k], hi = s.a_nodes[2 * k + 1];
        const float tx0 = (lo.x - o.x) * inv.x, tx1 = (hi.x - o.x) * inv.x;
        const float ty0 = (lo.y - o.y) * inv.y, ty1 = (hi.y - o.y) * inv.y;
        const float tz0 = (lo.z - o.z) * inv.z, tz1 = (hi.z - o.z) * inv.z;
        const float tnear = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), tmin));
        const float tfar = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
        const int skip = __float_as_int(lo.w);
        if (tnear > tfar || tnear > best) {
            k = skip;
            continue;
        }
        const uint32_t leaf = __float_as_uint(hi.w);
        const int count = (int)(leaf >> 28);
        if (count == 0) {
            ++k;
            continue;
        }
        const int first = (int)(leaf & 0x0FFFFFFFu);
        for (int j = first; j < first + count; ++j) {
            const int i = s.a_order[j];
            float t;
            f3 p, n;
            if (!alight_hit(s.analytic[i], o, d, tmin, tmax, &t, &p, &n)) continue;
            if (t < best || (t == best && i < found)) {
                best = t;
                found = i;
            }
        }
        k = skip;
    }
    return found;
}

// analytic_light_pdf_for_hit (programs.cu:815-834)
__device__ __forceinline__ float alight_pdf_for_hit(const rt_analytic_light& L, f3 origin, f3 dir,
    f3 hit_pos, f3 hit_n) {
    if (L.delta != 0 || L.treat_as_point != 0 || L.treat_as_line != 0) return 0.0f;
    if (L.type == RT_ALIGHT_SPHERE) {
        const float radius = L.radius * len3(ld3(L.basis_x));
        const float dsq = lensq3(ld3(L.position) - origin);
        const float rsq = radius * radius;
        if (dsq <= rsq) return 0.0f;
        return cone_pdf(sqrtf(fmaxf(0.0f, 1.0f - rsq / dsq)));
    }
    const float dsq = lensq3(hit_pos - origin);
    const float ac = fmaxf(dot3(hit_n, dir * -1.0f), 0.0f);
    return area_pdf(L.world_area, dsq, ac);
}

// analytic_emission_mis_weight (programs.cu:836-849): the light's hit along
// the path's ray (ro, rd) is recomputed; the pdf uses the normalized ray
__device__ __forceinline__ float analytic_mis(const SceneView& s, const PathPrev& pv, int li, f3 ro,
    f3 rd) {
    if (!pv.valid || pv.delta || li < 0) return 1.0f;
    const rt_analytic_light L = s.analytic[li];
    float t = 0.0f;
    f3 p = mk3(0, 0, 0), n = mk3(0, 0, 0);
    if (L.type != RT_ALIGHT_SPHERE) (void)alight_hit(L, ro, rd, kRayEps, kRayFar, &t, &p, &n);
    const float lp = L.selection_pdf * alight_pdf_for_hit(L, pv.pos, normalize3(rd), p, n);
    return power_heuristic(pv.pdf, lp);
}

// analytic_infinite_direction / analytic_infinite_pdf /
// analytic_infinite_radiance (programs.cu:851-890): the dome and distant
// lights seen along a missed ray, in index order
__device__ inline f3 analytic_infinite(const SceneView& s, const PathPrev& pv, f3 rd) {
    f3 sum = mk3(0, 0, 0);
    const f3 dir = normalize3(rd);
    for (int k = 0; k < s.a_infinite_count; ++k) {
        const rt_analytic_light L = s.analytic[s.a_infinite[k]];
        if (L.type == RT_ALIGHT_DISTANT) {
            const f3 axis = normalize3(ld3(L.basis_z));
            const float threshold = L.delta != 0 ? 0.999999f : L.cos_theta_max;
            if (!(dot3(axis, dir) >= threshold)) continue;
        }
        float w = 1.0f;
        if (pv.valid && !pv.delta) {
            const float cond = L.delta != 0 ? 0.0f
                               : (L.type == RT_ALIGHT_DOME ? kUniformSpherePdf : cone_pdf(L.cos_theta_max));
            if (cond <= 0.0f) continue;
            w = power_heuristic(pv.pdf, L.selection_pdf * cond);
        }
        sum = sum + ld3(L.radiance) * w;
    }
    return sum;
}

// ---- homogeneous media
// hit_medium_boundary (programs.cu:1356-1419): the ray's interval inside the
// medium's sphere or box, in the medium's frame, clipped to (tmin, tmax)
__device__ inline bool medium_boundary(const rt_medium& m, f3 o, f3 d, float tmin, float tmax,
    float* entry, float* exit) {
    const f3 r0 = ld3(m.rotation_row0), r1 = ld3(m.rotation_row1), r2 = ld3(m.rotation_row2);
    const f3 sh = o - ld3(m.translation);
    const f3 lo = mk3(dot3(r0, sh), dot3(r1, sh), dot3(r2, sh));
    const f3 ldir = mk3(dot3(r0, d), dot3(r1, d), dot3(r2, d));
    if (m.boundary_type == RT_MEDIUM_SPHERE) {
        const f3 oc = lo - ld3(m.local_center_or_min);
        const float a = lensq3(ldir);
        const float hb = dot3(oc, ldir);
        const float c = lensq3(oc) - m.radius * m.radius;
        const float disc = hb * hb - a * c;
        if (disc <= 0.0f) return false;
        const float sq = sqrtf(disc);
        float t0 = (-hb - sq) / a, t1 = (-hb + sq) / a;
        if (t0 > t1) {
            const float tt = t0;
            t0 = t1;
            t1 = tt;
        }
        *entry = fmaxf(t0, tmin);
        *exit = fminf(t1, tmax);
        return *entry < *exit;
    }
    if (m.boundary_type != RT_MEDIUM_BOX) return false;
    float t0 = tmin, t1 = tmax;
    const float org[3] = {lo.x, lo.y, lo.z}, dir[3] = {ldir.x, ldir.y, ldir.z};
    for (int a = 0; a < 3; ++a) {
        if (fabsf(dir[a]) < 1e-8f) {
            if (org[a] < m.local_center_or_min[a] || org[a] > m.local_max[a]) return false;
            continue;
        }
        float a0 = (m.local_center_or_min[a] - org[a]) / dir[a];
        float a1 = (m.local_max[a] - org[a]) / dir[a];
        if (a0 > a1) {
            const float tt = a0;
            a0 = a1;
            a1 = tt;
        }
        t0 = fmaxf(t0, a0);
        t1 = fminf(t1, a1);
        if (t0 >= t1) return false;
    }
    *entry = t0;
    *exit = t1;
    return true;
}

// the media part of intersect_scene (programs.cu:1711-1753): in index
// order, a medium the ray crosses before `closest` scatters it at distance
// -log(max(u, 1e-6)) / density into the medium if that falls inside;
// medium i draws u from the path's medium group 48 + i / 4, component i % 4.
// Returns the scattering medium (closest updated) or -1.
__device__ inline int sample_media(const SceneView& s, f3 o, f3 d, float tmin, float* closest,
    uint32_t pixel, uint32_t sample, uint32_t key, uint32_t bounce, uint32_t stream) {
    const float ray_len = len3(d);
    if (!(ray_len > 1e-8f)) return -1;
    int event = -1;
    float4 u4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (int i = 0; i < s.medium_count; ++i) {
        if ((i & 3) == 0)
            u4 = rng4(pixel, sample, key, bounce, kGroupMedium0 + (uint32_t)(i >> 2), stream);
        const rt_medium m = s.media[i];
        if (m.density <= 0.0f) continue;
        float entry = 0.0f, exit = 0.0f;
        if (!medium_boundary(m, o, d, tmin, *closest, &entry, &exit)) continue;
        entry = fmaxf(entry, tmin);
        exit = fminf(exit, *closest);
        if (entry >= exit) continue;
        const float inside = (exit - entry) * ray_len;
        const float u = (i & 3) == 0 ? u4.x : ((i & 3) == 1 ? u4.y : ((i & 3) == 2 ? u4.z : u4.w));
        const float dist = -log_fixed(fmaxf(u, 1e-6f)) / m.density;
        if (dist >= inside) continue;
        *closest = entry + dist / ray_len;
        event = i;
    }
    return event;
}

// sample_packed_analytic_light (light_sampling.h:104-116) by bisection
__device__ __forceinline__ int pick_analytic(const SceneView& s, float u) {
    if (s.analytic_count <= 0) return -1;
    const float x = u < 0.0f ? 0.0f : (u < 1.0f ? u : 0.99999994f);
    int lo = 0, hi = s.analytic_count - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (x < s.a_cdf[mid]) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// sample_analytic_direct_light_candidate (programs.cu:1234-1346)
__device__ inline LightSample analytic_candidate(const SceneView& s, f3 sp, int li, float u0,
    float u1) {
    LightSample ls{};
    ls.valid = false;
    ls.infinite = false;
    ls.delta = false;
    if (li < 0 || li >= s.analytic_count) return ls;
    const rt_analytic_light L = s.analytic[li];
    if (L.selection_pdf <= 0.0f) return ls;
    const f3 emission = ld3(L.radiance);
    if (L.type == RT_ALIGHT_DOME) {
        ls.dir = sphere_dir(u0, u1);
        ls.dist = kRayFar;
        ls.emission = emission;
        ls.pdf = L.selection_pdf * kUniformSpherePdf;
        ls.infinite = true;
        ls.valid = ls.pdf > 0.0f && maxc3(ls.emission) > 0.0f;
        return ls;
    }
    if (L.type == RT_ALIGHT_DISTANT) {
        const f3 axis = normalize3(ld3(L.basis_z));
        ls.dir = L.delta != 0 ? axis : cone_dir(axis, L.cos_theta_max, u0, u1);
        ls.dist = kRayFar;
        ls.emission = emission;
        ls.pdf = L.selection_pdf * (L.delta != 0 ? 1.0f : cone_pdf(L.cos_theta_max));
        ls.infinite = true;
        ls.delta = L.delta != 0;
        ls.valid = ls.pdf > 0.0f && maxc3(ls.emission) > 0.0f;
        return ls;
    }
    const f3 center = ld3(L.position);
    if (L.type == RT_ALIGHT_SPHERE) {
        const f3 tc = center - sp;
        const float dsq = lensq3(tc);
        const float radius = L.radius * len3(ld3(L.basis_x));
        const float rsq = radius * radius;
        if (dsq <= rsq * (1.0f + 1e-6f)) return ls;
        if (L.treat_as_point != 0) {
            ls.dist = sqrtf(dsq);
            ls.dir = div3(tc, ls.dist);
            const float scale = L.world_area > 1e-12f ? L.world_area : 1.0f;
            ls.emission = emission * (scale / dsq);
            ls.pdf = L.selection_pdf;
            ls.delta = true;
            ls.valid = maxc3(ls.emission) > 0.0f;
            return ls;
        }
        const float cmax = sqrtf(fmaxf(0.0f, 1.0f - rsq / dsq));
        const f3 dir = cone_dir(normalize3(tc), cmax, u0, u1);
        float t;
        f3 p, n;
        if (!alight_hit(L, sp, dir, kRayEps, kRayFar, &t, &p, &n)) return ls;
        ls.dir = dir;
        ls.emission = emission;
        ls.dist = t;
        ls.pdf = L.selection_pdf * cone_pdf(cmax);
        ls.valid = ls.pdf > 0.0f && maxc3(ls.emission) > 0.0f;
        return ls;
    }
    f3 pos, nrm;
    if (L.type == RT_ALIGHT_DISK) {
        const float radial = L.radius * sqrtf(u0);
        float sphi, cphi;
        sincos_turn(u1, &sphi, &cphi);
        pos = center + (ld3(L.basis_x) * (radial * cphi) + ld3(L.basis_y) * (radial * sphi));
        nrm = alight_normal(L);
    } else if (L.type == RT_ALIGHT_RECT) {
        pos = center + (ld3(L.basis_x) * ((u0 - 0.5f) * L.width) + ld3(L.basis_y) * ((u1 - 0.5f) * L.height));
        nrm = alight_normal(L);
    } else {
        float sphi, cphi;
        sincos_turn(u0, &sphi, &cphi);
        const float rx = L.radius * cphi, ry = L.radius * sphi;
        pos = center + ((ld3(L.basis_x) * rx + ld3(L.basis_y) * ry) + ld3(L.basis_z) * ((u1 - 0.5f) * L.length));
        nrm = normalize3(ld3(L.basis_x) * rx + ld3(L.basis_y) * ry);
    }
    const f3 tl = pos - sp;
    const float dsq = lensq3(tl);
    if (dsq <= 1e-12f) return ls;
    ls.dist = sqrtf(dsq);
    ls.dir = div3(tl, ls.dist);
    const float ac = dot3(nrm, ls.dir * -1.0f);
    ls.emission = emission;
    ls.pdf = L.selection_pdf * area_pdf(L.world_area, dsq, ac);
    ls.valid = ls.pdf > 0.0f && maxc3(ls.emission) > 0.0f;
    return ls;
}

__device__ __forceinline__ float schlick(float cosine, float ri) {
    float r0 = (1.0f - ri) / (1.0f + ri);
    r0 *= r0;
    const float x = 1.0f - cosine;
    const float x2 = x * x;
    return r0 + (1.0f - r0) * (x2 * x2 * x);
}

struct BsdfResult {
    f3 o, d, weight;
    float pdf;
    bool delta, alive;
    int event;   // OpenPBR scatter event (kPbr*), 0 for the other materials
};

template <int PBR = kPbrFull, bool PTEX = false>
__device__ inline BsdfResult sample_bsdf(const SceneView& s, const HitInfo& h, f3 ray_d, float4 r,
    f3 thr) {
    BsdfResult b{};
    b.alive = false;
    b.event = 0;
    if (h.mat_type == RT_MAT_DIFFUSE_LIGHT) return b;
    if (h.mat_type == RT_MAT_LAMBERTIAN) {
        f3 sc = h.sn + sphere_dir(r.x, r.y);
        if (near_zero3(sc)) sc = h.sn;
        b.o = h.pos + h.sn * kRayEps;
        b.d = normalize3(sc);
        b.weight = h.base;
        b.pdf = fmaxf(dot3(h.sn, b.d), 0.0f) / kPi;
        b.delta = false;
        b.alive = true;
        return b;
    }
    if (h.mat_type == RT_MAT_METAL) {
        const f3 refl = reflect3(normalize3(ray_d), h.sn);
        const f3 ball = sphere_dir(r.x, r.y) * cbrt01(r.z);
        const f3 fz = ball * fminf(fmaxf(h.fuzz, 0.0f), 1.0f);
        const f3 sc = normalize3(refl + fz);
        if (dot3(sc, h.sn) <= 0.0f) return b;
        b.o = h.pos + h.sn * kRayEps;
        b.d = sc;
        b.weight = h.base;
        b.pdf = 0.0f;
        b.delta = true;
        b.alive = true;
        return b;
    }
    if (h.mat_type == RT_MAT_DIELECTRIC) {
        const float ratio = h.front ? (1.0f / fmaxf(h.ior, 1e-4f)) : fmaxf(h.ior, 1e-4f);
        const f3 unit = normalize3(ray_d);
        const float ct = fminf(dot3(-unit, h.sn), 1.0f);
        const float stt = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
        const bool cannot = ratio * stt > 1.0f;
        f3 dir;
        if (cannot || schlick(ct, ratio) > r.x) dir = reflect3(unit, h.sn);
        else dir = refract3(unit, h.sn, ratio);
        b.o = h.pos + dir * kRayEps;
        b.d = normalize3(dir);
        b.weight = h.base;
        b.pdf = 0.0f;
        b.delta = true;
        b.alive = true;
        return b;
    }
    if (PBR != kPbrOff && h.mat_type == RT_MAT_OPENPBR) {
        OpenPbrSampleResult smp;
        const rt_openpbr_material* pm = s.openpbr + h.openpbr;
        auto sample = [&](const auto& m) {
            openpbr_sample_v(m, h.gn, -normalize3(ray_d), r.x, r.y, r.z, thr, &smp);
        };
        if constexpr (PBR == kPbrCore) with_pbr_material<PTEX>(s, opbr::MatCore(pm), h, sample);
        else with_pbr_material<PTEX>(s, opbr::MatMem{pm}, h, sample);
        if (!smp.valid) return b;
        b.event = smp.event;
        b.o = h.pos + smp.wi * kRayEps;
        b.d = normalize3(smp.wi);
        b.weight = smp.weight;
        b.pdf = smp.pdf;
        b.delta = smp.delta;
        b.alive = true;
        return b;
    }
    if (h.mat_type == RT_MAT_ISOTROPIC) {
        const f3 dir = sphere_dir(r.x, r.y);
        b.o = h.pos + dir * kRayEps;
        b.d = dir;
        b.weight = h.base;
        b.pdf = kUniformSpherePdf;
        b.delta = false;
        b.alive = true;
        return b;
    }
    return b;
}

}  // namespace rtk
