// rt_openpbr.h -- OpenPBR surface with constant inputs for the shade stage
// (and for the host light-CDF builder, hence __host__ __device__).
//
// Follows src/common/openpbr_core.h for the lobes a constant-input surface
// without transmission, subsurface, fuzz, thin film, dispersion, anisotropy
// or thin walls uses (such materials are rejected at scene upload):
//   Fujii Oren-Nayar base diffuse with multi-scatter       :1389-1459
//   GGX dielectric / F82 metal specular with energy comp.   :915-1016,1490-1677
//   GGX coat (delta when smooth), passage and darkening     :1044-1107,1211-1298
//   lobe probabilities :1306-1387, evaluate :1679-1860,
//   emission :1868-1883, sample :1995-2297
// Terms of the omitted lobes vanish exactly for these inputs.  Sampled
// azimuths use sincos_turn instead of cosf/sinf(2*pi*u).
#pragma once

#include "rt_internal.h"
#include "rt_math.h"

namespace rtk {

#define RTK_HD __host__ __device__ __forceinline__

struct OpenPbrSampleResult {
    f3 wi, weight;
    float pdf;
    bool delta, valid;
    int event;
};

namespace opbr {

constexpr float kPiF = 3.14159265358979323846f;
constexpr float kInvPi = 1.0f / kPiF;

struct Rough {
    float ax, ay;
};
struct Frame {
    f3 t, b, n;
};
struct Probs {
    float diffuse, reflection, coat;
};

RTK_HD f3 v(float x, float y, float z) { return f3{x, y, z}; }
RTK_HD f3 vs(float s) { return f3{s, s, s}; }
RTK_HD f3 add(f3 a, f3 b) { return v(a.x + b.x, a.y + b.y, a.z + b.z); }
RTK_HD f3 sub(f3 a, f3 b) { return v(a.x - b.x, a.y - b.y, a.z - b.z); }
RTK_HD f3 mul(f3 a, f3 b) { return v(a.x * b.x, a.y * b.y, a.z * b.z); }
RTK_HD f3 scl(f3 a, float s) { return v(a.x * s, a.y * s, a.z * s); }
RTK_HD f3 neg(f3 a) { return v(-a.x, -a.y, -a.z); }
RTK_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RTK_HD f3 cross(f3 a, f3 b) {
    return v(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
RTK_HD float lensq(f3 a) { return dot(a, a); }
RTK_HD f3 ld(const float* p) { return v(p[0], p[1], p[2]); }
RTK_HD float omin(float a, float b) { return a < b ? a : b; }
RTK_HD float omax(float a, float b) { return a > b ? a : b; }
RTK_HD float oclamp(float x, float lo, float hi) { return omin(omax(x, lo), hi); }
RTK_HD float sat(float x) { return oclamp(x, 0.0f, 1.0f); }
RTK_HD float safe_sqrt(float x) { return sqrtf(omax(0.0f, x)); }
RTK_HD f3 clamp_unit(f3 a) { return v(sat(a.x), sat(a.y), sat(a.z)); }
RTK_HD f3 clamp_nonneg(f3 a) { return v(omax(0.0f, a.x), omax(0.0f, a.y), omax(0.0f, a.z)); }
RTK_HD f3 odiv(f3 a, float s) {
    const float inv = 1.0f / s;
    return scl(a, inv);
}
RTK_HD f3 lerp(f3 a, f3 b, float t) { return add(scl(a, 1.0f - t), scl(b, t)); }
RTK_HD float lum(f3 a) {
    return 0.2126f * omax(0.0f, a.x) + 0.7152f * omax(0.0f, a.y) + 0.0722f * omax(0.0f, a.z);
}
RTK_HD f3 onormalize(f3 a) {
    const float l2 = lensq(a);
    if (l2 <= 1e-20f) return vs(0.0f);
    return scl(a, 1.0f / sqrtf(l2));
}
RTK_HD float pow5(float x) {
    const float s = x * x;
    return s * s * x;
}

RTK_HD Frame make_frame(f3 normal) {
    Frame f;
    f.n = onormalize(normal);
    if (lensq(f.n) <= 0.0f) f.n = v(0.0f, 0.0f, 1.0f);
    const f3 helper = fabsf(f.n.z) < 0.999f ? v(0.0f, 0.0f, 1.0f) : v(0.0f, 1.0f, 0.0f);
    f.t = onormalize(cross(helper, f.n));
    f.b = onormalize(cross(f.n, f.t));
    f.t = onormalize(cross(f.b, f.n));
    return f;
}
RTK_HD Frame face(Frame f, f3 wo) {
    if (dot(f.n, wo) < 0.0f) {
        f.n = neg(f.n);
        f.b = neg(f.b);
    }
    return f;
}
RTK_HD f3 to_local(const Frame& f, f3 a) { return v(dot(a, f.t), dot(a, f.b), dot(a, f.n)); }
RTK_HD f3 to_world(const Frame& f, f3 a) {
    return add(add(scl(f.t, a.x), scl(f.b, a.y)), scl(f.n, a.z));
}

RTK_HD Rough roughness(float rv) {
    const float r = sat(rv);
    const float aniso = sat(0.0f);
    const float aspect = 1.0f - aniso;
    const float scale = sqrtf(2.0f / omax(aspect * aspect + 1.0f, 1e-8f));
    const float ax = omax(r * r * scale, 1e-4f);
    return Rough{ax, omax(aspect * ax, 1e-4f)};
}
RTK_HD float eff_spec_roughness(const rt_openpbr_material& m) {
    const float s = sat(m.specular_roughness);
    const float c = sat(m.coat_roughness);
    const float s2 = s * s, c2 = c * c;
    const float affected = sqrtf(sqrtf(omin(1.0f, s2 * s2 + 2.0f * c2 * c2)));
    const float w = sat(m.coat_weight);
    return s + (affected - s) * w;
}
RTK_HD Rough spec_rough(const rt_openpbr_material& m) { return roughness(eff_spec_roughness(m)); }
RTK_HD Rough coat_rough(const rt_openpbr_material& m) { return roughness(m.coat_roughness); }
RTK_HD bool smooth(Rough r) { return omax(r.ax, r.ay) < 1e-3f; }
RTK_HD float avg_alpha(Rough r) { return sqrtf(r.ax * r.ay); }

RTK_HD float ggx_albedo(float cosine, float alpha, float c0, float c90) {
    const float x = sat(cosine), y = sat(alpha);
    const float x2 = x * x, y2 = y * y;
    const float na = 0.1003f - 0.6303f * x + 9.748f * y - 2.038f * x * y + 29.34f * x2
                     - 8.245f * y2 - 26.44f * x2 * y + 19.99f * x * y2 - 5.448f * x2 * y2;
    const float nb = 0.9345f - 2.323f * x + 2.229f * y - 3.748f * x * y + 1.424f * x2
                     - 0.7684f * y2 + 1.436f * x2 * y + 0.2913f * x * y2 + 0.6286f * x2 * y2;
    const float da = 1.0f - 1.765f * x + 8.263f * y + 11.53f * x * y + 28.96f * x2
                     - 7.507f * y2 - 36.11f * x2 * y + 15.86f * x * y2 + 33.37f * x2 * y2;
    const float db = 1.0f + 0.2281f * x + 15.94f * y - 55.83f * x * y + 13.08f * x2
                     + 41.26f * y2 + 54.9f * x2 * y + 300.2f * x * y2 - 285.1f * x2 * y2;
    const float ca = sat(na / omax(da, 1e-6f));
    const float cb = sat(nb / omax(db, 1e-6f));
    return sat(c0) * ca + sat(c90) * cb;
}
RTK_HD f3 energy_comp(float cosine, float alpha, f3 f) {
    if (alpha < 1e-3f) return vs(1.0f);
    const float ss = ggx_albedo(cosine, alpha, 1.0f, 1.0f);
    const float missing = (1.0f - ss) / omax(ss, 1e-4f);
    return add(vs(1.0f), scl(clamp_unit(f), missing));
}
RTK_HD float ggx_d(Rough r, f3 wm) {
    if (wm.z <= 0.0f) return 0.0f;
    const float x = wm.x / r.ax, y = wm.y / r.ay;
    const float den = x * x + y * y + wm.z * wm.z;
    return 1.0f / (kPiF * r.ax * r.ay * den * den);
}
RTK_HD float ggx_lambda(Rough r, f3 a) {
    const float az = fabsf(a.z);
    if (az <= 1e-8f) return 1e20f;
    const float proj = r.ax * r.ax * a.x * a.x + r.ay * r.ay * a.y * a.y;
    const float at2 = proj / (az * az);
    return 0.5f * (sqrtf(1.0f + at2) - 1.0f);
}
RTK_HD float ggx_g1(Rough r, f3 a) { return 1.0f / (1.0f + ggx_lambda(r, a)); }
RTK_HD float ggx_g(Rough r, f3 wo, f3 wi) {
    return 1.0f / (1.0f + ggx_lambda(r, wo) + ggx_lambda(r, wi));
}
RTK_HD float vndf_pdf(Rough r, f3 wo, f3 wm) {
    const float aco = fabsf(wo.z);
    if (aco <= 1e-8f) return 0.0f;
    return ggx_g1(r, wo) * ggx_d(r, wm) * fabsf(dot(wo, wm)) / aco;
}
RTK_HD f3 sample_vndf(Rough r, f3 wo, float u1, float u2) {
    f3 wh = onormalize(v(r.ax * wo.x, r.ay * wo.y, wo.z));
    if (wh.z < 0.0f) wh = neg(wh);
    const f3 t1 = wh.z < 0.99999f ? onormalize(cross(v(0.0f, 0.0f, 1.0f), wh)) : v(1.0f, 0.0f, 0.0f);
    const f3 t2 = cross(wh, t1);
    const float radius = sqrtf(sat(u1));
    float sp, cp;
    sincos_turn(sat(u2), &sp, &cp);
    const float dx = radius * cp;
    float dy = radius * sp;
    const float horizon = safe_sqrt(1.0f - dx * dx);
    const float warp = 0.5f * (1.0f + wh.z);
    dy = horizon * (1.0f - warp) + dy * warp;
    const float dz = safe_sqrt(1.0f - dx * dx - dy * dy);
    const f3 nh = add(add(scl(t1, dx), scl(t2, dy)), scl(wh, dz));
    return onormalize(v(r.ax * nh.x, r.ay * nh.y, omax(1e-6f, nh.z)));
}
RTK_HD f3 reflect(f3 wo, f3 wm) { return sub(scl(wm, 2.0f * dot(wo, wm)), wo); }

RTK_HD float fresnel_dielectric(float cosine, float eta) {
    cosine = oclamp(cosine, -1.0f, 1.0f);
    eta = omax(eta, 1e-4f);
    if (cosine < 0.0f) {
        eta = 1.0f / eta;
        cosine = -cosine;
    }
    const float s2i = omax(0.0f, 1.0f - cosine * cosine);
    const float s2t = s2i / (eta * eta);
    if (s2t >= 1.0f) return 1.0f;
    const float ct = safe_sqrt(1.0f - s2t);
    const float par = (eta * cosine - ct) / omax(eta * cosine + ct, 1e-8f);
    const float perp = (cosine - eta * ct) / omax(cosine + eta * ct, 1e-8f);
    return 0.5f * (par * par + perp * perp);
}
RTK_HD float modulated_ior(const rt_openpbr_material& m) {
    const float eta = omax(m.specular_ior, 1e-4f);
    const float ratio = (eta - 1.0f) / (eta + 1.0f);
    const float f0 = oclamp(omax(0.0f, m.specular_weight) * ratio * ratio, 0.0f, 0.9999f);
    const float sign = eta < 1.0f ? -1.0f : (eta > 1.0f ? 1.0f : 0.0f);
    const float eps = sign * sqrtf(f0);
    return omax((1.0f + eps) / omax(1.0f - eps, 1e-4f), 1e-4f);
}
RTK_HD float dielectric_fresnel(const rt_openpbr_material& m, float cosine, bool entering) {
    float eta = modulated_ior(m);
    if (!entering) eta = 1.0f / eta;
    return sat(fresnel_dielectric(cosine, eta));
}
RTK_HD f3 schlick(f3 c0, f3 c90, float cosine) { return lerp(c0, c90, pow5(1.0f - sat(cosine))); }
RTK_HD f3 metal_fresnel(const rt_openpbr_material& m, float cosine) {
    const float kEdge = 1.0f / 7.0f;
    const float mu = sat(fabsf(cosine));
    const f3 c0 = scl(clamp_unit(ld(m.base_color)), sat(m.base_weight));
    const f3 c90 = vs(1.0f);
    const f3 sch = schlick(c0, c90, mu);
    const f3 sch_edge = schlick(c0, c90, kEdge);
    const f3 desired = mul(sch_edge, clamp_unit(ld(m.specular_color)));
    const float omm = 1.0f - mu;
    const float ome = 1.0f - kEdge;
    const float corr = mu * pow5(omm) * omm / (kEdge * pow5(ome) * ome);
    const f3 f82 = sub(sch, scl(sub(sch_edge, desired), corr));
    return scl(clamp_unit(f82), sat(m.specular_weight));
}
RTK_HD f3 dielectric_albedo_rgb(const rt_openpbr_material& m, float cosine, bool entering) {
    const Rough r = spec_rough(m);
    const float alpha = avg_alpha(r);
    if (smooth(r)) return vs(dielectric_fresnel(m, cosine, entering));
    const float c0 = dielectric_fresnel(m, 1.0f, entering);
    if (c0 <= 1e-8f) return vs(0.0f);
    const float ss = ggx_albedo(cosine, alpha, c0, 1.0f);
    const float white = ggx_albedo(cosine, alpha, 1.0f, 1.0f);
    const float df = dielectric_fresnel(m, cosine, entering);
    const float comp = 1.0f + df * (1.0f - white) / omax(white, 1e-4f);
    return vs(sat(ss * comp));
}
RTK_HD f3 reflection_fresnel(const rt_openpbr_material& m, float mcos, float ocos, bool entering) {
    const float metal = sat(m.base_metalness);
    const float alpha = avg_alpha(spec_rough(m));
    const f3 df = vs(dielectric_fresnel(m, mcos, entering));
    const f3 dsingle = mul(entering ? clamp_unit(ld(m.specular_color)) : vs(1.0f), df);
    const f3 dielectric = mul(dsingle, energy_comp(ocos, alpha, df));
    const f3 msingle = metal_fresnel(m, mcos);
    const f3 metal_v = mul(msingle, energy_comp(ocos, alpha, msingle));
    return lerp(dielectric, metal_v, metal);
}

RTK_HD float coat_fresnel(const rt_openpbr_material& m, float cosine) {
    return fresnel_dielectric(cosine, omax(m.coat_ior, 1e-4f));
}
RTK_HD float coat_albedo(const rt_openpbr_material& m, float cosine) {
    const Rough r = coat_rough(m);
    if (smooth(r)) return coat_fresnel(m, cosine);
    const float alpha = avg_alpha(r);
    const float c0 = coat_fresnel(m, 1.0f);
    const float ss = ggx_albedo(cosine, alpha, c0, 1.0f);
    const float white = ggx_albedo(cosine, alpha, 1.0f, 1.0f);
    const float comp = 1.0f + coat_fresnel(m, cosine) * (1.0f - white) / omax(white, 1e-4f);
    return sat(ss * comp);
}
RTK_HD f3 coat_passage(const rt_openpbr_material& m, float cosine) {
    const float presence = sat(m.coat_weight);
    const f3 tint = clamp_unit(ld(m.coat_color));
    if (presence <= 0.0f || cosine <= 0.0f || (tint.x >= 1.0f && tint.y >= 1.0f && tint.z >= 1.0f))
        return vs(1.0f);
    const float eta = omax(m.coat_ior, 1e-4f);
    const float s2t = omax(0.0f, 1.0f - cosine * cosine) / (eta * eta);
    const float ct = safe_sqrt(1.0f - omin(s2t, 1.0f));
    const float ds = 1.0f / omax(ct, 1e-4f);
    const f3 st = v(safe_sqrt(tint.x), safe_sqrt(tint.y), safe_sqrt(tint.z));
    const f3 pw = v(powf(omax(st.x, 0.0f), ds), powf(omax(st.y, 0.0f), ds), powf(omax(st.z, 0.0f), ds));
    return lerp(vs(1.0f), pw, presence);
}
RTK_HD f3 coat_darkening(const rt_openpbr_material& m) {
    const float presence = sat(m.coat_weight) * sat(m.coat_darkening);
    if (presence <= 0.0f) return vs(1.0f);
    const float eta = omax(m.coat_ior, 1e-4f);
    const float f0 = coat_fresnel(m, 1.0f);
    const float ir = sat(1.0f - (1.0f - f0) / omax(eta * eta, 1e-6f));
    const f3 da = scl(clamp_unit(ld(m.base_color)), sat(m.base_weight));
    const f3 ma = scl(da, sat(m.specular_weight));
    const f3 ba = lerp(da, ma, sat(m.base_metalness));
    const float num = 1.0f - ir;
    const f3 dk = v(num / omax(1.0f - ba.x * ir, 1e-5f), num / omax(1.0f - ba.y * ir, 1e-5f),
        num / omax(1.0f - ba.z * ir, 1e-5f));
    return lerp(vs(1.0f), clamp_unit(dk), presence);
}
RTK_HD f3 substrate_scale(const rt_openpbr_material& m, float ocos, float icos, bool exterior) {
    const float presence = sat(m.coat_weight);
    if (presence <= 0.0f) return vs(1.0f);
    if (!exterior) return icos < 0.0f ? coat_passage(m, -icos) : vs(1.0f);
    const float rout = presence * coat_albedo(m, ocos);
    const f3 out = scl(coat_passage(m, ocos), sat(1.0f - rout));
    f3 in = vs(1.0f);
    if (icos > 0.0f) {
        const float rin = presence * coat_albedo(m, icos);
        in = scl(coat_passage(m, icos), sat(1.0f - rin));
    }
    return mul(mul(out, in), coat_darkening(m));
}
RTK_HD f3 coat_value(const rt_openpbr_material& m, f3 wo, f3 wi) {
    const float presence = sat(m.coat_weight);
    if (presence <= 0.0f || wo.z <= 0.0f || wi.z <= 0.0f) return vs(0.0f);
    const Rough r = coat_rough(m);
    if (smooth(r)) return vs(0.0f);
    const f3 sm = add(wi, wo);
    if (lensq(sm) <= 1e-16f) return vs(0.0f);
    f3 wm = onormalize(sm);
    if (wm.z < 0.0f) wm = neg(wm);
    const float wodm = dot(wo, wm);
    if (wodm <= 0.0f) return vs(0.0f);
    const float D = ggx_d(r, wm);
    const float G = ggx_g(r, wo, wi);
    const float F = coat_fresnel(m, wodm);
    const float comp = energy_comp(wo.z, avg_alpha(r), vs(F)).x;
    return vs(presence * F * comp * D * G / omax(4.0f * wo.z * wi.z, 1e-8f));
}
RTK_HD float coat_pdf(const rt_openpbr_material& m, f3 wo, f3 wi) {
    if (wo.z <= 0.0f || wi.z <= 0.0f) return 0.0f;
    const Rough r = coat_rough(m);
    if (smooth(r)) return 0.0f;
    const f3 sm = add(wi, wo);
    if (lensq(sm) <= 1e-16f) return 0.0f;
    f3 wm = onormalize(sm);
    if (wm.z < 0.0f) wm = neg(wm);
    const float wodm = dot(wo, wm);
    if (wodm <= 0.0f) return 0.0f;
    return vndf_pdf(r, wo, wm) / omax(4.0f * fabsf(wodm), 1e-8f);
}
RTK_HD Probs lobe_probs(const rt_openpbr_material& m, float ocos, bool exterior) {
    Probs p{0.0f, 0.0f, 0.0f};
    const float metal = sat(m.base_metalness);
    const float bw = sat(m.base_weight);
    const float msw = sat(m.specular_weight);
    const float eta = modulated_ior(m);
    const float er = (eta - 1.0f) / (eta + 1.0f);
    const float df0 = er * er;
    const float msr = msw * lum(scl(clamp_unit(ld(m.base_color)), bw));
    const float opaque = (1.0f - metal) * (1.0f - 0.0f) * sat(1.0f - df0);
    const float diffuse = opaque * bw * lum(clamp_unit(ld(m.base_color)));
    const float drs = (1.0f - metal) * df0 * lum(clamp_unit(ld(m.specular_color)));
    const float mrs = metal * msr;
    const float coat = exterior ? sat(m.coat_weight) * coat_albedo(m, ocos) : 0.0f;
    const float sss = sat(1.0f - coat) * sat(1.0f - 0.0f);
    const float sd = diffuse * sss;
    const float rs = (drs + mrs) * sss;
    const float total = sd + rs + coat;
    if (total <= 1e-10f) return p;
    p.diffuse = sd / total;
    p.reflection = rs / total;
    p.coat = coat / total;
    return p;
}
RTK_HD float on_dir_albedo(float cosine, float roughness_v) {
    const float k1 = 0.5f - 2.0f / (3.0f * kPiF);
    const float c = oclamp(cosine, 1e-6f, 1.0f);
    const float sigma = sat(roughness_v);
    const float a = 1.0f / (1.0f + k1 * sigma);
    const float b = sigma * a;
    const float s = safe_sqrt(1.0f - c * c);
    const float g = s * (acosf(c) - s * c) + 2.0f * ((s / c) * (1.0f - s * s * s) - s) / 3.0f;
    return sat(a + b * g * kInvPi);
}
RTK_HD float on_avg_albedo(float roughness_v) {
    const float k1 = 0.5f - 2.0f / (3.0f * kPiF);
    const float k2 = 2.0f / 3.0f - 28.0f / (15.0f * kPiF);
    const float sigma = sat(roughness_v);
    const float a = 1.0f / (1.0f + k1 * sigma);
    return sat(a * (1.0f + k2 * sigma));
}
RTK_HD f3 diffuse(const rt_openpbr_material& m, f3 wo, f3 wi) {
    if (wo.z <= 0.0f || wi.z <= 0.0f) return vs(0.0f);
    const float metal = sat(m.base_metalness);
    const float mix = sat(m.base_weight);
    const f3 color = clamp_unit(ld(m.base_color));
    const float coeff = mix * (1.0f - metal) * (1.0f - 0.0f);
    if (coeff <= 0.0f) return vs(0.0f);
    const float k1 = 0.5f - 2.0f / (3.0f * kPiF);
    const float sigma = sat(m.base_diffuse_roughness);
    const float a = 1.0f / (1.0f + k1 * sigma);
    const float s = dot(wi, wo) - wi.z * wo.z;
    const float sot = s > 0.0f ? s / omax(wi.z, wo.z) : s;
    const f3 single = scl(color, a * (1.0f + sigma * sot));
    const float din = on_dir_albedo(wi.z, sigma);
    const float dout = on_dir_albedo(wo.z, sigma);
    const float avg = on_avg_albedo(sigma);
    const float miss = omax(1.0f - avg, 1e-6f);
    const f3 c2 = mul(color, color);
    const f3 msc = v(c2.x * avg / omax(1.0f - color.x * (1.0f - avg), 1e-6f),
        c2.y * avg / omax(1.0f - color.y * (1.0f - avg), 1e-6f),
        c2.z * avg / omax(1.0f - color.z * (1.0f - avg), 1e-6f));
    const float mss = (1.0f - din) * (1.0f - dout) / miss;
    const f3 multi = scl(msc, mss);
    return scl(add(single, multi), coeff * kInvPi);
}

}  // namespace opbr

RTK_HD f3 openpbr_emission(const rt_openpbr_material& m) {
    using namespace opbr;
    f3 r = scl(clamp_nonneg(ld(m.emission_color)), omax(0.0f, m.emission_luminance) * sat(m.geometry_opacity));
    const float ct = 1.0f - sat(m.coat_weight) * coat_albedo(m, 1.0f);
    r = mul(r, scl(coat_passage(m, 1.0f), sat(ct)));
    return r;
}

// evaluate with frame = make_openpbr_frame(n, 0); wo/wi world-space
RTK_HD void openpbr_eval(const rt_openpbr_material& m, f3 n, f3 wo_world, f3 wi_world, f3* value,
    float* pdf) {
    using namespace opbr;
    *value = vs(0.0f);
    *pdf = 0.0f;
    const f3 won = onormalize(wo_world);
    const f3 win = onormalize(wi_world);
    if (lensq(won) <= 0.0f || lensq(win) <= 0.0f) return;
    Frame f = make_frame(n);
    const bool entering = dot(f.n, won) >= 0.0f;
    f = face(f, won);
    const f3 wo = to_local(f, won);
    const f3 wi = to_local(f, win);
    if (wo.z <= 1e-7f || fabsf(wi.z) <= 1e-7f) return;
    const float opacity = sat(m.geometry_opacity);
    const bool exterior = entering;
    const Probs p = lobe_probs(m, wo.z, exterior);
    const Rough r = spec_rough(m);
    const bool is_smooth = smooth(r);
    f3 val = vs(0.0f);
    float pd = 0.0f;
    if (wi.z * wo.z > 0.0f) {
        const f3 att = sub(vs(1.0f), dielectric_albedo_rgb(m, wo.z, entering));
        val = diffuse(m, wo, wi);
        val = mul(val, clamp_unit(att));
        pd = p.diffuse * wi.z * kInvPi;
        const f3 sm = add(wi, wo);
        if (!is_smooth && lensq(sm) > 1e-16f) {
            f3 wm = onormalize(sm);
            if (wm.z < 0.0f) wm = neg(wm);
            const float wodm = dot(wo, wm);
            if (wodm > 0.0f) {
                const float D = ggx_d(r, wm);
                const float G = ggx_g(r, wo, wi);
                const f3 F = reflection_fresnel(m, wodm, wo.z, entering);
                const f3 spec = scl(F, D * G / omax(4.0f * wo.z * wi.z, 1e-8f));
                val = add(val, spec);
                const float mpdf = vndf_pdf(r, wo, wm) / omax(4.0f * fabsf(wodm), 1e-8f);
                pd += p.reflection * mpdf;
            }
        }
        val = mul(val, substrate_scale(m, wo.z, wi.z, exterior));
        if (exterior) {
            val = add(val, coat_value(m, wo, wi));
            pd += p.coat * coat_pdf(m, wo, wi);
        }
    }
    *value = scl(val, opacity);
    *pdf = pd * opacity;
}

RTK_HD void openpbr_sample(const rt_openpbr_material& m, f3 n, f3 wo_world, float u_lobe, float u1,
    float u2, OpenPbrSampleResult* out) {
    using namespace opbr;
    out->wi = vs(0.0f);
    out->weight = vs(0.0f);
    out->pdf = 0.0f;
    out->delta = false;
    out->valid = false;
    out->event = 0;
    const f3 won = onormalize(wo_world);
    if (lensq(won) <= 0.0f) return;
    Frame f = make_frame(n);
    const bool entering = dot(f.n, won) >= 0.0f;
    f = face(f, won);
    const f3 wo = to_local(f, won);
    if (wo.z <= 1e-7f) return;
    const float opacity = sat(m.geometry_opacity);
    const float pass = 1.0f - opacity;
    u_lobe = sat(u_lobe);
    if (u_lobe < pass) {
        out->wi = neg(won);
        out->weight = vs(1.0f);
        out->event = 5;
        out->delta = true;
        out->valid = true;
        return;
    }
    if (opacity <= 0.0f) return;
    const float choice = oclamp((u_lobe - pass) / opacity, 0.0f, 0.99999994f);
    const bool exterior = entering;
    const Probs p = lobe_probs(m, wo.z, exterior);
    if (p.diffuse + p.reflection + p.coat <= 0.0f) return;
    const float coat_end = p.coat;
    const float diffuse_end = coat_end + p.diffuse;
    const float reflection_end = diffuse_end + p.reflection;
    f3 wi;
    int event;
    if (choice < coat_end) {
        const Rough cr = coat_rough(m);
        if (smooth(cr)) {
            wi = v(-wo.x, -wo.y, wo.z);
            const f3 coeff = vs(sat(m.coat_weight) * coat_fresnel(m, wo.z) * opacity);
            const f3 w = onormalize(to_world(f, wi));
            const float dp = opacity * p.coat;
            if (dp <= 0.0f) return;
            out->wi = w;
            out->weight = odiv(coeff, dp);
            out->event = 6;
            out->delta = true;
            out->valid = true;
            return;
        }
        const f3 wm = sample_vndf(cr, wo, u1, u2);
        wi = reflect(wo, wm);
        if (wi.z <= 1e-7f) return;
        event = 6;
    } else if (choice < diffuse_end) {
        const float radius = sqrtf(sat(u1));
        float sp, cp;
        sincos_turn(sat(u2), &sp, &cp);
        wi = v(radius * cp, radius * sp, safe_sqrt(1.0f - radius * radius));
        event = 1;
    } else if (choice < reflection_end) {
        const Rough r = spec_rough(m);
        if (smooth(r)) {
            wi = v(-wo.x, -wo.y, wo.z);
            f3 coeff = scl(reflection_fresnel(m, wo.z, wo.z, entering), opacity);
            coeff = mul(coeff, substrate_scale(m, wo.z, wi.z, exterior));
            const f3 w = onormalize(to_world(f, wi));
            const float dp = opacity * p.reflection;
            if (dp <= 0.0f) return;
            out->wi = w;
            out->weight = odiv(coeff, dp);
            out->event = 2;
            out->delta = true;
            out->valid = true;
            return;
        }
        const f3 wm = sample_vndf(r, wo, u1, u2);
        wi = reflect(wo, wm);
        if (wi.z <= 1e-7f) return;
        event = 2;
    } else {
        return;  // zero-probability transmission lobe: the reference yields an invalid sample
    }
    const f3 w = onormalize(to_world(f, wi));
    f3 val;
    float pdf;
    openpbr_eval(m, n, won, w, &val, &pdf);
    if (pdf <= 1e-12f) return;
    out->wi = w;
    out->weight = scl(val, fabsf(dot(f.n, w)) / pdf);
    out->pdf = pdf;
    out->event = event;
    out->delta = false;
    out->valid = true;
}

#undef RTK_HD

}  // namespace rtk
