/*
 * oracle_openpbr.c -- OpenPBR surface, constant-input subset.
 * TEST INFRASTRUCTURE ONLY (see rt_oracle.h).
 *
 * Restates src/common/openpbr_core.h for materials whose transmission,
 * subsurface, fuzz and thin-film weights, dispersion scale, anisotropies and
 * thin-walled flag are zero (the product rejects other materials at upload):
 *   roughness / effective specular roughness  :863-912
 *   GGX albedo fit, energy compensation, D, G, VNDF :915-1016
 *   dielectric Fresnel :592-610, modulated IOR :815-823
 *   coat Fresnel / albedo / passage / darkening :1044-1107
 *   coat base scale, coat reflection value/pdf :1211-1298
 *   lobe probabilities :1306-1387
 *   Fujii Oren-Nayar diffuse + multi-scatter :1389-1459
 *   F82 metal Fresnel :1490-1508, reflection Fresnel :1655-1677
 *   evaluate :1679-1860, emission :1868-1883, sample :1995-2297
 * Fuzz, transmission and subsurface terms vanish exactly for these inputs
 * and are omitted.  The azimuth of every sampled direction uses the own
 * sin/cos of a turn (om.h) instead of cosf/sinf(2*pi*u).
 */
#include "oracle_internal.h"

#define OP_PI 3.14159265358979323846f
#define OP_INV_PI (1.0f / OP_PI)

typedef struct { float ax, ay; } op_rough;
typedef struct { v3 t, b, n; } op_frame;

static float op_min(float a, float b) { return a < b ? a : b; }
static float op_max(float a, float b) { return a > b ? a : b; }
static float op_clamp(float v, float lo, float hi) { return op_min(op_max(v, lo), hi); }
static float op_sat(float v) { return op_clamp(v, 0.0f, 1.0f); }
static float op_safe_sqrt(float v) { return sqrtf(op_max(0.0f, v)); }
static v3 op_vec(float v) { return mk3(v, v, v); }
static v3 op_clamp_unit(v3 v) { return mk3(op_sat(v.x), op_sat(v.y), op_sat(v.z)); }
static v3 op_clamp_nonneg(v3 v) { return mk3(op_max(0.0f, v.x), op_max(0.0f, v.y), op_max(0.0f, v.z)); }
static v3 op_div(v3 v, float s) { const float inv = 1.0f / s; return scl3(v, inv); }
static v3 op_lerp(v3 a, v3 b, float t) { return add3(scl3(a, 1.0f - t), scl3(b, t)); }
static float op_lum(v3 v) {
    return 0.2126f * op_max(0.0f, v.x) + 0.7152f * op_max(0.0f, v.y) + 0.0722f * op_max(0.0f, v.z);
}
static v3 op_normalize(v3 v) {
    const float l2 = lensq3(v);
    if (l2 <= 1e-20f) return mk3(0.0f, 0.0f, 0.0f);
    return scl3(v, 1.0f / sqrtf(l2));
}
static float op_pow5(float v) { const float s = v * v; return s * s * v; }

static op_frame op_make_frame(v3 normal) {
    op_frame f;
    f.n = op_normalize(normal);
    if (lensq3(f.n) <= 0.0f) f.n = mk3(0.0f, 0.0f, 1.0f);
    /* the authored tangent is zero, so the helper-axis branch is taken */
    const v3 helper = fabsf(f.n.z) < 0.999f ? mk3(0.0f, 0.0f, 1.0f) : mk3(0.0f, 1.0f, 0.0f);
    f.t = op_normalize(cross3(helper, f.n));
    f.b = op_normalize(cross3(f.n, f.t));
    f.t = op_normalize(cross3(f.b, f.n));
    return f;
}
static op_frame op_face(op_frame f, v3 wo) {
    if (dot3(f.n, wo) < 0.0f) { f.n = neg3(f.n); f.b = neg3(f.b); }
    return f;
}
static v3 op_to_local(const op_frame* f, v3 v) { return mk3(dot3(v, f->t), dot3(v, f->b), dot3(v, f->n)); }
static v3 op_to_world(const op_frame* f, v3 v) {
    return add3(add3(scl3(f->t, v.x), scl3(f->b, v.y)), scl3(f->n, v.z));
}

static op_rough op_roughness(float rv) {
    const float r = op_sat(rv);
    const float aniso = op_sat(0.0f);
    const float aspect = 1.0f - aniso;
    const float scale = sqrtf(2.0f / op_max(aspect * aspect + 1.0f, 1e-8f));
    const float ax = op_max(r * r * scale, 1e-4f);
    op_rough o = {ax, op_max(aspect * ax, 1e-4f)};
    return o;
}
static float op_eff_spec_roughness(const rto_openpbr* m) {
    const float s = op_sat(m->specular_roughness);
    const float c = op_sat(m->coat_roughness);
    const float s2 = s * s, c2 = c * c;
    const float affected = sqrtf(sqrtf(op_min(1.0f, s2 * s2 + 2.0f * c2 * c2)));
    const float w = op_sat(m->coat_weight);
    return s + (affected - s) * w;
}
static op_rough op_spec_rough(const rto_openpbr* m) { return op_roughness(op_eff_spec_roughness(m)); }
static op_rough op_coat_rough(const rto_openpbr* m) { return op_roughness(m->coat_roughness); }
static int op_smooth(op_rough r) { return op_max(r.ax, r.ay) < 1e-3f; }
static float op_avg_alpha(op_rough r) { return sqrtf(r.ax * r.ay); }

static float op_ggx_albedo(float cosine, float alpha, float c0, float c90) {
    const float x = op_sat(cosine), y = op_sat(alpha);
    const float x2 = x * x, y2 = y * y;
    const float na = 0.1003f - 0.6303f * x + 9.748f * y - 2.038f * x * y + 29.34f * x2
                     - 8.245f * y2 - 26.44f * x2 * y + 19.99f * x * y2 - 5.448f * x2 * y2;
    const float nb = 0.9345f - 2.323f * x + 2.229f * y - 3.748f * x * y + 1.424f * x2
                     - 0.7684f * y2 + 1.436f * x2 * y + 0.2913f * x * y2 + 0.6286f * x2 * y2;
    const float da = 1.0f - 1.765f * x + 8.263f * y + 11.53f * x * y + 28.96f * x2
                     - 7.507f * y2 - 36.11f * x2 * y + 15.86f * x * y2 + 33.37f * x2 * y2;
    const float db = 1.0f + 0.2281f * x + 15.94f * y - 55.83f * x * y + 13.08f * x2
                     + 41.26f * y2 + 54.9f * x2 * y + 300.2f * x * y2 - 285.1f * x2 * y2;
    const float ca = op_sat(na / op_max(da, 1e-6f));
    const float cb = op_sat(nb / op_max(db, 1e-6f));
    return op_sat(c0) * ca + op_sat(c90) * cb;
}
static v3 op_energy_comp(float cosine, float alpha, v3 f) {
    if (alpha < 1e-3f) return op_vec(1.0f);
    const float ss = op_ggx_albedo(cosine, alpha, 1.0f, 1.0f);
    const float missing = (1.0f - ss) / op_max(ss, 1e-4f);
    return add3(op_vec(1.0f), scl3(op_clamp_unit(f), missing));
}
static float op_ggx_d(op_rough r, v3 wm) {
    if (wm.z <= 0.0f) return 0.0f;
    const float x = wm.x / r.ax, y = wm.y / r.ay;
    const float den = x * x + y * y + wm.z * wm.z;
    return 1.0f / (OP_PI * r.ax * r.ay * den * den);
}
static float op_ggx_lambda(op_rough r, v3 v) {
    const float az = fabsf(v.z);
    if (az <= 1e-8f) return 1e20f;
    const float proj = r.ax * r.ax * v.x * v.x + r.ay * r.ay * v.y * v.y;
    const float at2 = proj / (az * az);
    return 0.5f * (sqrtf(1.0f + at2) - 1.0f);
}
static float op_ggx_g1(op_rough r, v3 v) { return 1.0f / (1.0f + op_ggx_lambda(r, v)); }
static float op_ggx_g(op_rough r, v3 wo, v3 wi) {
    return 1.0f / (1.0f + op_ggx_lambda(r, wo) + op_ggx_lambda(r, wi));
}
static float op_vndf_pdf(op_rough r, v3 wo, v3 wm) {
    const float aco = fabsf(wo.z);
    if (aco <= 1e-8f) return 0.0f;
    return op_ggx_g1(r, wo) * op_ggx_d(r, wm) * fabsf(dot3(wo, wm)) / aco;
}
static v3 op_sample_vndf(op_rough r, v3 wo, float u1, float u2) {
    v3 wh = op_normalize(mk3(r.ax * wo.x, r.ay * wo.y, wo.z));
    if (wh.z < 0.0f) wh = neg3(wh);
    const v3 t1 = wh.z < 0.99999f ? op_normalize(cross3(mk3(0.0f, 0.0f, 1.0f), wh)) : mk3(1.0f, 0.0f, 0.0f);
    const v3 t2 = cross3(wh, t1);
    const float radius = sqrtf(op_sat(u1));
    float sp, cp;
    om_sincos_turn(op_sat(u2), &sp, &cp);
    const float dx = radius * cp;
    float dy = radius * sp;
    const float horizon = op_safe_sqrt(1.0f - dx * dx);
    const float warp = 0.5f * (1.0f + wh.z);
    dy = horizon * (1.0f - warp) + dy * warp;
    const float dz = op_safe_sqrt(1.0f - dx * dx - dy * dy);
    const v3 nh = add3(add3(scl3(t1, dx), scl3(t2, dy)), scl3(wh, dz));
    return op_normalize(mk3(r.ax * nh.x, r.ay * nh.y, op_max(1e-6f, nh.z)));
}
static v3 op_reflect(v3 wo, v3 wm) { return sub3(scl3(wm, 2.0f * dot3(wo, wm)), wo); }

static float op_fresnel_dielectric(float cosine, float eta) {
    cosine = op_clamp(cosine, -1.0f, 1.0f);
    eta = op_max(eta, 1e-4f);
    if (cosine < 0.0f) { eta = 1.0f / eta; cosine = -cosine; }
    const float s2i = op_max(0.0f, 1.0f - cosine * cosine);
    const float s2t = s2i / (eta * eta);
    if (s2t >= 1.0f) return 1.0f;
    const float ct = op_safe_sqrt(1.0f - s2t);
    const float par = (eta * cosine - ct) / op_max(eta * cosine + ct, 1e-8f);
    const float perp = (cosine - eta * ct) / op_max(cosine + eta * ct, 1e-8f);
    return 0.5f * (par * par + perp * perp);
}
static float op_modulated_ior(const rto_openpbr* m) {
    const float eta = op_max(m->specular_ior, 1e-4f);
    const float ratio = (eta - 1.0f) / (eta + 1.0f);
    const float f0 = op_clamp(op_max(0.0f, m->specular_weight) * ratio * ratio, 0.0f, 0.9999f);
    const float sign = eta < 1.0f ? -1.0f : (eta > 1.0f ? 1.0f : 0.0f);
    const float eps = sign * sqrtf(f0);
    return op_max((1.0f + eps) / op_max(1.0f - eps, 1e-4f), 1e-4f);
}
static float op_dielectric_fresnel(const rto_openpbr* m, float cosine, int entering) {
    float eta = op_modulated_ior(m);
    if (!entering) eta = 1.0f / eta;
    return op_sat(op_fresnel_dielectric(cosine, eta));
}
static v3 op_schlick(v3 c0, v3 c90, float cosine) {
    return op_lerp(c0, c90, op_pow5(1.0f - op_sat(cosine)));
}
static v3 op_metal_fresnel(const rto_openpbr* m, float cosine) {
    const float kEdge = 1.0f / 7.0f;
    const float mu = op_sat(fabsf(cosine));
    const v3 c0 = scl3(op_clamp_unit(ld3(m->base_color)), op_sat(m->base_weight));
    const v3 c90 = op_vec(1.0f);
    const v3 sch = op_schlick(c0, c90, mu);
    const v3 sch_edge = op_schlick(c0, c90, kEdge);
    const v3 desired = mul3(sch_edge, op_clamp_unit(ld3(m->specular_color)));
    const float omm = 1.0f - mu;
    const float ome = 1.0f - kEdge;
    const float corr = mu * op_pow5(omm) * omm / (kEdge * op_pow5(ome) * ome);
    const v3 f82 = sub3(sch, scl3(sub3(sch_edge, desired), corr));
    return scl3(op_clamp_unit(f82), op_sat(m->specular_weight));
}
static v3 op_dielectric_albedo_rgb(const rto_openpbr* m, float cosine, int entering) {
    const op_rough r = op_spec_rough(m);
    const float alpha = op_avg_alpha(r);
    if (op_smooth(r)) return op_vec(op_dielectric_fresnel(m, cosine, entering));
    const float c0 = op_dielectric_fresnel(m, 1.0f, entering);
    if (c0 <= 1e-8f) return op_vec(0.0f);
    const float ss = op_ggx_albedo(cosine, alpha, c0, 1.0f);
    const float white = op_ggx_albedo(cosine, alpha, 1.0f, 1.0f);
    const float df = op_dielectric_fresnel(m, cosine, entering);
    const float comp = 1.0f + df * (1.0f - white) / op_max(white, 1e-4f);
    return op_vec(op_sat(ss * comp));
}
static v3 op_reflection_fresnel(const rto_openpbr* m, float mcos, float ocos, int entering) {
    const float metal = op_sat(m->base_metalness);
    const float alpha = op_avg_alpha(op_spec_rough(m));
    const v3 df = op_vec(op_dielectric_fresnel(m, mcos, entering));
    const v3 dsingle = mul3(entering ? op_clamp_unit(ld3(m->specular_color)) : op_vec(1.0f), df);
    const v3 dielectric = mul3(dsingle, op_energy_comp(ocos, alpha, df));
    const v3 msingle = op_metal_fresnel(m, mcos);
    const v3 metal_v = mul3(msingle, op_energy_comp(ocos, alpha, msingle));
    return op_lerp(dielectric, metal_v, metal);
}

static float op_coat_fresnel(const rto_openpbr* m, float cosine) {
    return op_fresnel_dielectric(cosine, op_max(m->coat_ior, 1e-4f));
}
static float op_coat_albedo(const rto_openpbr* m, float cosine) {
    const op_rough r = op_coat_rough(m);
    if (op_smooth(r)) return op_coat_fresnel(m, cosine);
    const float alpha = op_avg_alpha(r);
    const float c0 = op_coat_fresnel(m, 1.0f);
    const float ss = op_ggx_albedo(cosine, alpha, c0, 1.0f);
    const float white = op_ggx_albedo(cosine, alpha, 1.0f, 1.0f);
    const float comp = 1.0f + op_coat_fresnel(m, cosine) * (1.0f - white) / op_max(white, 1e-4f);
    return op_sat(ss * comp);
}
static v3 op_coat_passage(const rto_openpbr* m, float cosine) {
    const float presence = op_sat(m->coat_weight);
    const v3 tint = op_clamp_unit(ld3(m->coat_color));
    if (presence <= 0.0f || cosine <= 0.0f || (tint.x >= 1.0f && tint.y >= 1.0f && tint.z >= 1.0f))
        return op_vec(1.0f);
    const float eta = op_max(m->coat_ior, 1e-4f);
    const float s2t = op_max(0.0f, 1.0f - cosine * cosine) / (eta * eta);
    const float ct = op_safe_sqrt(1.0f - op_min(s2t, 1.0f));
    const float ds = 1.0f / op_max(ct, 1e-4f);
    const v3 st = mk3(op_safe_sqrt(tint.x), op_safe_sqrt(tint.y), op_safe_sqrt(tint.z));
    const v3 pw = mk3(powf(op_max(st.x, 0.0f), ds), powf(op_max(st.y, 0.0f), ds), powf(op_max(st.z, 0.0f), ds));
    return op_lerp(op_vec(1.0f), pw, presence);
}
static v3 op_coat_darkening(const rto_openpbr* m) {
    const float presence = op_sat(m->coat_weight) * op_sat(m->coat_darkening);
    if (presence <= 0.0f) return op_vec(1.0f);
    const float eta = op_max(m->coat_ior, 1e-4f);
    const float f0 = op_coat_fresnel(m, 1.0f);
    const float ir = op_sat(1.0f - (1.0f - f0) / op_max(eta * eta, 1e-6f));
    const v3 da = scl3(op_clamp_unit(ld3(m->base_color)), op_sat(m->base_weight));
    const v3 ma = scl3(da, op_sat(m->specular_weight));
    const v3 ba = op_lerp(da, ma, op_sat(m->base_metalness));
    const float num = 1.0f - ir;
    const v3 dk = mk3(num / op_max(1.0f - ba.x * ir, 1e-5f), num / op_max(1.0f - ba.y * ir, 1e-5f),
        num / op_max(1.0f - ba.z * ir, 1e-5f));
    return op_lerp(op_vec(1.0f), op_clamp_unit(dk), presence);
}
static v3 op_substrate_scale(const rto_openpbr* m, float ocos, float icos, int exterior) {
    const float presence = op_sat(m->coat_weight);
    if (presence <= 0.0f) return op_vec(1.0f);
    if (!exterior) return icos < 0.0f ? op_coat_passage(m, -icos) : op_vec(1.0f);
    const float rout = presence * op_coat_albedo(m, ocos);
    const v3 out = scl3(op_coat_passage(m, ocos), op_sat(1.0f - rout));
    v3 in = op_vec(1.0f);
    if (icos > 0.0f) {
        const float rin = presence * op_coat_albedo(m, icos);
        in = scl3(op_coat_passage(m, icos), op_sat(1.0f - rin));
    }
    return mul3(mul3(out, in), op_coat_darkening(m));
}
static v3 op_coat_value(const rto_openpbr* m, v3 wo, v3 wi) {
    const float presence = op_sat(m->coat_weight);
    if (presence <= 0.0f || wo.z <= 0.0f || wi.z <= 0.0f) return op_vec(0.0f);
    const op_rough r = op_coat_rough(m);
    if (op_smooth(r)) return op_vec(0.0f);
    const v3 sum = add3(wi, wo);
    if (lensq3(sum) <= 1e-16f) return op_vec(0.0f);
    v3 wm = op_normalize(sum);
    if (wm.z < 0.0f) wm = neg3(wm);
    const float wodm = dot3(wo, wm);
    if (wodm <= 0.0f) return op_vec(0.0f);
    const float D = op_ggx_d(r, wm);
    const float G = op_ggx_g(r, wo, wi);
    const float F = op_coat_fresnel(m, wodm);
    const float comp = op_energy_comp(wo.z, op_avg_alpha(r), op_vec(F)).x;
    return op_vec(presence * F * comp * D * G / op_max(4.0f * wo.z * wi.z, 1e-8f));
}
static float op_coat_pdf(const rto_openpbr* m, v3 wo, v3 wi) {
    if (wo.z <= 0.0f || wi.z <= 0.0f) return 0.0f;
    const op_rough r = op_coat_rough(m);
    if (op_smooth(r)) return 0.0f;
    const v3 sum = add3(wi, wo);
    if (lensq3(sum) <= 1e-16f) return 0.0f;
    v3 wm = op_normalize(sum);
    if (wm.z < 0.0f) wm = neg3(wm);
    const float wodm = dot3(wo, wm);
    if (wodm <= 0.0f) return 0.0f;
    return op_vndf_pdf(r, wo, wm) / op_max(4.0f * fabsf(wodm), 1e-8f);
}

typedef struct { float diffuse, reflection, coat; } op_probs;

static op_probs op_lobe_probs(const rto_openpbr* m, float ocos, int exterior) {
    op_probs p = {0.0f, 0.0f, 0.0f};
    const float metal = op_sat(m->base_metalness);
    const float bw = op_sat(m->base_weight);
    const float msw = op_sat(m->specular_weight);
    const float eta = op_modulated_ior(m);
    const float er = (eta - 1.0f) / (eta + 1.0f);
    const float df0 = er * er;
    const float msr = msw * op_lum(scl3(op_clamp_unit(ld3(m->base_color)), bw));
    const float opaque = (1.0f - metal) * (1.0f - 0.0f) * op_sat(1.0f - df0);
    const float diffuse = opaque * bw * op_lum(op_clamp_unit(ld3(m->base_color)));
    const float drs = (1.0f - metal) * df0 * op_lum(op_clamp_unit(ld3(m->specular_color)));
    const float mrs = metal * msr;
    const float coat = exterior ? op_sat(m->coat_weight) * op_coat_albedo(m, ocos) : 0.0f;
    const float sss = op_sat(1.0f - coat) * op_sat(1.0f - 0.0f);
    const float sd = diffuse * sss;
    const float rs = (drs + mrs) * sss;
    const float total = sd + rs + coat;
    if (total <= 1e-10f) return p;
    p.diffuse = sd / total;
    p.reflection = rs / total;
    p.coat = coat / total;
    return p;
}

static float op_on_dir_albedo(float cosine, float roughness) {
    const float k1 = 0.5f - 2.0f / (3.0f * OP_PI);
    const float c = op_clamp(cosine, 1e-6f, 1.0f);
    const float sigma = op_sat(roughness);
    const float a = 1.0f / (1.0f + k1 * sigma);
    const float b = sigma * a;
    const float s = op_safe_sqrt(1.0f - c * c);
    const float g = s * (acosf(c) - s * c) + 2.0f * ((s / c) * (1.0f - s * s * s) - s) / 3.0f;
    return op_sat(a + b * g * OP_INV_PI);
}
static float op_on_avg_albedo(float roughness) {
    const float k1 = 0.5f - 2.0f / (3.0f * OP_PI);
    const float k2 = 2.0f / 3.0f - 28.0f / (15.0f * OP_PI);
    const float sigma = op_sat(roughness);
    const float a = 1.0f / (1.0f + k1 * sigma);
    return op_sat(a * (1.0f + k2 * sigma));
}
static v3 op_diffuse(const rto_openpbr* m, v3 wo, v3 wi) {
    if (wo.z <= 0.0f || wi.z <= 0.0f) return op_vec(0.0f);
    const float metal = op_sat(m->base_metalness);
    const float mix = op_sat(m->base_weight);
    const v3 color = op_clamp_unit(ld3(m->base_color));
    const float coeff = mix * (1.0f - metal) * (1.0f - 0.0f);
    if (coeff <= 0.0f) return op_vec(0.0f);
    const float k1 = 0.5f - 2.0f / (3.0f * OP_PI);
    const float sigma = op_sat(m->base_diffuse_roughness);
    const float a = 1.0f / (1.0f + k1 * sigma);
    const float s = dot3(wi, wo) - wi.z * wo.z;
    const float sot = s > 0.0f ? s / op_max(wi.z, wo.z) : s;
    const v3 single = scl3(color, a * (1.0f + sigma * sot));
    const float din = op_on_dir_albedo(wi.z, sigma);
    const float dout = op_on_dir_albedo(wo.z, sigma);
    const float avg = op_on_avg_albedo(sigma);
    const float miss = op_max(1.0f - avg, 1e-6f);
    const v3 c2 = mul3(color, color);
    const v3 msc = mk3(c2.x * avg / op_max(1.0f - color.x * (1.0f - avg), 1e-6f),
        c2.y * avg / op_max(1.0f - color.y * (1.0f - avg), 1e-6f),
        c2.z * avg / op_max(1.0f - color.z * (1.0f - avg), 1e-6f));
    const float mss = (1.0f - din) * (1.0f - dout) / miss;
    const v3 multi = scl3(msc, mss);
    return scl3(add3(single, multi), coeff * OP_INV_PI);
}

v3 om_openpbr_emission(const rto_openpbr* m) {
    v3 r = scl3(op_clamp_nonneg(ld3(m->emission_color)),
        op_max(0.0f, m->emission_luminance) * op_sat(m->geometry_opacity));
    const float ct = 1.0f - op_sat(m->coat_weight) * op_coat_albedo(m, 1.0f);
    r = mul3(r, scl3(op_coat_passage(m, 1.0f), op_sat(ct)));
    return r;
}

/* evaluate_openpbr_core with frame = make_openpbr_frame(n, 0) */
void om_openpbr_eval(const rto_openpbr* m, v3 n, v3 wo_world, v3 wi_world, v3* value, float* pdf) {
    *value = op_vec(0.0f);
    *pdf = 0.0f;
    const v3 won = op_normalize(wo_world);
    const v3 win = op_normalize(wi_world);
    if (lensq3(won) <= 0.0f || lensq3(win) <= 0.0f) return;
    op_frame f = op_make_frame(n);
    const int entering = dot3(f.n, won) >= 0.0f;
    f = op_face(f, won);
    const v3 wo = op_to_local(&f, won);
    const v3 wi = op_to_local(&f, win);
    if (wo.z <= 1e-7f || fabsf(wi.z) <= 1e-7f) return;
    const float opacity = op_sat(m->geometry_opacity);
    const int exterior = entering;
    const op_probs p = op_lobe_probs(m, wo.z, exterior);
    const op_rough r = op_spec_rough(m);
    const int smooth = op_smooth(r);
    v3 val = op_vec(0.0f);
    float pd = 0.0f;
    if (wi.z * wo.z > 0.0f) {
        const v3 att = sub3(op_vec(1.0f), op_dielectric_albedo_rgb(m, wo.z, entering));
        val = op_diffuse(m, wo, wi);
        val = mul3(val, op_clamp_unit(att));
        pd = p.diffuse * wi.z * OP_INV_PI;
        const v3 sum = add3(wi, wo);
        if (!smooth && lensq3(sum) > 1e-16f) {
            v3 wm = op_normalize(sum);
            if (wm.z < 0.0f) wm = neg3(wm);
            const float wodm = dot3(wo, wm);
            if (wodm > 0.0f) {
                const float D = op_ggx_d(r, wm);
                const float G = op_ggx_g(r, wo, wi);
                const v3 F = op_reflection_fresnel(m, wodm, wo.z, entering);
                const v3 spec = scl3(F, D * G / op_max(4.0f * wo.z * wi.z, 1e-8f));
                val = add3(val, spec);
                const float mpdf = op_vndf_pdf(r, wo, wm) / op_max(4.0f * fabsf(wodm), 1e-8f);
                pd += p.reflection * mpdf;
            }
        }
        val = mul3(val, op_substrate_scale(m, wo.z, wi.z, exterior));
        if (exterior) {
            val = add3(val, op_coat_value(m, wo, wi));
            pd += p.coat * op_coat_pdf(m, wo, wi);
        }
    }
    *value = scl3(val, opacity);
    *pdf = pd * opacity;
}

/* sample_openpbr_core with frame = make_openpbr_frame(n, 0) */
void om_openpbr_sample(const rto_openpbr* m, v3 n, v3 wo_world, float u_lobe, float u1, float u2,
    rto_openpbr_sample* out) {
    memset(out, 0, sizeof(*out));
    const v3 won = op_normalize(wo_world);
    if (lensq3(won) <= 0.0f) return;
    op_frame f = op_make_frame(n);
    const int entering = dot3(f.n, won) >= 0.0f;
    f = op_face(f, won);
    const v3 wo = op_to_local(&f, won);
    if (wo.z <= 1e-7f) return;
    const float opacity = op_sat(m->geometry_opacity);
    const float pass = 1.0f - opacity;
    u_lobe = op_sat(u_lobe);
    if (u_lobe < pass) {
        const v3 wi = neg3(won);
        out->wi[0] = wi.x; out->wi[1] = wi.y; out->wi[2] = wi.z;
        out->weight[0] = out->weight[1] = out->weight[2] = 1.0f;
        out->event = 5;
        out->delta = 1;
        out->valid = 1;
        return;
    }
    if (opacity <= 0.0f) return;
    const float choice = op_clamp((u_lobe - pass) / opacity, 0.0f, 0.99999994f);
    const int exterior = entering;
    const op_probs p = op_lobe_probs(m, wo.z, exterior);
    if (p.diffuse + p.reflection + p.coat <= 0.0f) return;
    const float coat_end = p.coat;
    const float diffuse_end = coat_end + p.diffuse;
    const float reflection_end = diffuse_end + p.reflection;
    v3 wi;
    int event;
    if (choice < coat_end) {
        const op_rough cr = op_coat_rough(m);
        if (op_smooth(cr)) {
            wi = mk3(-wo.x, -wo.y, wo.z);
            const v3 coeff = op_vec(op_sat(m->coat_weight) * op_coat_fresnel(m, wo.z) * opacity);
            const v3 w = op_normalize(op_to_world(&f, wi));
            const float dp = opacity * p.coat;
            if (dp <= 0.0f) return;
            const v3 weight = op_div(coeff, dp);
            out->wi[0] = w.x; out->wi[1] = w.y; out->wi[2] = w.z;
            out->weight[0] = weight.x; out->weight[1] = weight.y; out->weight[2] = weight.z;
            out->event = 6;
            out->delta = 1;
            out->valid = 1;
            return;
        }
        const v3 wm = op_sample_vndf(cr, wo, u1, u2);
        wi = op_reflect(wo, wm);
        if (wi.z <= 1e-7f) return;
        event = 6;
    } else if (choice < diffuse_end) {
        const float radius = sqrtf(op_sat(u1));
        float sp, cp;
        om_sincos_turn(op_sat(u2), &sp, &cp);
        wi = mk3(radius * cp, radius * sp, op_safe_sqrt(1.0f - radius * radius));
        event = 1;
    } else if (choice < reflection_end) {
        const op_rough r = op_spec_rough(m);
        if (op_smooth(r)) {
            wi = mk3(-wo.x, -wo.y, wo.z);
            v3 coeff = scl3(op_reflection_fresnel(m, wo.z, wo.z, entering), opacity);
            coeff = mul3(coeff, op_substrate_scale(m, wo.z, wi.z, exterior));
            const v3 w = op_normalize(op_to_world(&f, wi));
            const float dp = opacity * p.reflection;
            if (dp <= 0.0f) return;
            const v3 weight = op_div(coeff, dp);
            out->wi[0] = w.x; out->wi[1] = w.y; out->wi[2] = w.z;
            out->weight[0] = weight.x; out->weight[1] = weight.y; out->weight[2] = weight.z;
            out->event = 2;
            out->delta = 1;
            out->valid = 1;
            return;
        }
        const v3 wm = op_sample_vndf(r, wo, u1, u2);
        wi = op_reflect(wo, wm);
        if (wi.z <= 1e-7f) return;
        event = 2;
    } else {
        /* transmission lobe has zero probability for these materials: every
         * reference branch beyond reflection_end yields an invalid sample */
        return;
    }
    const v3 w = op_normalize(op_to_world(&f, wi));
    v3 val;
    float pdf;
    om_openpbr_eval(m, n, won, w, &val, &pdf);
    if (pdf <= 1e-12f) return;
    const v3 weight = scl3(val, fabsf(dot3(f.n, w)) / pdf);
    out->wi[0] = w.x; out->wi[1] = w.y; out->wi[2] = w.z;
    out->weight[0] = weight.x; out->weight[1] = weight.y; out->weight[2] = weight.z;
    out->pdf = pdf;
    out->event = event;
    out->delta = 0;
    out->valid = 1;
}

void rto_openpbr_evaluate(const rto_openpbr* m, const float n[3], const float wo[3],
    const float wi[3], rto_openpbr_eval* out) {
    v3 val;
    om_openpbr_eval(m, ld3(n), ld3(wo), ld3(wi), &val, &out->pdf);
    out->value[0] = val.x; out->value[1] = val.y; out->value[2] = val.z;
}

void rto_openpbr_sample_fn(const rto_openpbr* m, const float n[3], const float wo[3], float u_lobe,
    float u1, float u2, rto_openpbr_sample* out) {
    om_openpbr_sample(m, ld3(n), ld3(wo), u_lobe, u1, u2, out);
}

void rto_openpbr_emission(const rto_openpbr* m, float out[3]) {
    const v3 e = om_openpbr_emission(m);
    out[0] = e.x; out[1] = e.y; out[2] = e.z;
}
