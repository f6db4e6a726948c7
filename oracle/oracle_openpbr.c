/*
 * oracle_openpbr.c -- the OpenPBR surface (src/common/openpbr_core.h).
 * TEST INFRASTRUCTURE ONLY (see rt_oracle.h).
 *
 * Restates every lobe of the reference's OpenPBR core:
 *   transport context (RGB wavelengths, path throughput) :27-31,255-270
 *   transmission colour :360-366; subsurface medium, segment sampling and
 *     Henyey-Greenstein phase :387-519,576-590
 *   Fresnel (scalar, polarized, real-IOR phase) :592-660
 *   thin-film Airy reflectance (Belcour-Barla sensitivity) :662-813
 *   modulated IOR, dispersion :815-861
 *   anisotropic roughness, GGX albedo fit, energy compensation, D, G, VNDF,
 *     reflect / refract :863-1042
 *   coat Fresnel / albedo / passage / darkening :1044-1107
 *   fuzz (LTC sheen) :1109-1209; coat and fuzz base scale :1211-1242
 *   coat reflection value / pdf :1244-1298; lobe probabilities :1306-1387
 *   Fujii Oren-Nayar diffuse with multi-scatter, thin-walled subsurface
 *     :1389-1473
 *   dielectric / metal Fresnel, thin-wall window, thin film applied,
 *     dielectric and metal reflectance, directional albedo, reflection
 *     Fresnel :1475-1677
 *   evaluate :1679-1860, emission :1868-1883, subsurface boundary
 *     :1891-1993, sample :1995-2297
 * The azimuth of every sampled direction uses the own sin/cos of a turn
 * (om.h) instead of cosf/sinf(2*pi*u); the thin-film sensitivity keeps the
 * libm cosf / expf of its published fit.
 */
#include "oracle_internal.h"

#define OP_PI 3.14159265358979323846f
#define OP_INV_PI (1.0f / OP_PI)
#define OP_FRAUNHOFER_C 656.3f
#define OP_FRAUNHOFER_D 587.6f
#define OP_FRAUNHOFER_F 486.1f

typedef struct { float ax, ay; } op_rough;
typedef struct { v3 t, b, n; } op_frame;
typedef struct { float diffuse, subsurface, reflection, transmission, coat, fuzz; } op_probs;
typedef struct { float p, s; } op_pol;

/* ---- scalar and vector helpers */
static float op_min(float a, float b) { return a < b ? a : b; }
static float op_max(float a, float b) { return a > b ? a : b; }
static float op_clamp(float v, float lo, float hi) { return op_min(op_max(v, lo), hi); }
static float op_sat(float v) { return op_clamp(v, 0.0f, 1.0f); }
static float op_sq(float v) { return v * v; }
static float op_safe_sqrt(float v) { return sqrtf(op_max(0.0f, v)); }
static v3 op_vec(float v) { return mk3(v, v, v); }
static v3 op_clamp_unit(v3 v) { return mk3(op_sat(v.x), op_sat(v.y), op_sat(v.z)); }
static v3 op_clamp_nonneg(v3 v) { return mk3(op_max(0.0f, v.x), op_max(0.0f, v.y), op_max(0.0f, v.z)); }
static v3 op_div(v3 v, float s) { const float inv = 1.0f / s; return scl3(v, inv); }
static v3 op_lerp(v3 a, v3 b, float t) { return add3(scl3(a, 1.0f - t), scl3(b, t)); }
static v3 op_recip(v3 v) {
    return mk3(1.0f / op_max(v.x, 1e-6f), 1.0f / op_max(v.y, 1e-6f), 1.0f / op_max(v.z, 1e-6f));
}
static float op_comp(v3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
static void op_set_comp(v3* v, int i, float x) {
    if (i == 0) v->x = x; else if (i == 1) v->y = x; else v->z = x;
}
static float op_lum(v3 v) {
    return 0.2126f * op_max(0.0f, v.x) + 0.7152f * op_max(0.0f, v.y) + 0.0722f * op_max(0.0f, v.z);
}
static v3 op_normalize(v3 v) {
    const float l2 = lensq3(v);
    if (l2 <= 1e-20f) return mk3(0.0f, 0.0f, 0.0f);
    return scl3(v, 1.0f / sqrtf(l2));
}
static float op_pow5(float v) { const float s = v * v; return s * s * v; }
static v3 op_sqrt3(v3 v) { return mk3(op_safe_sqrt(v.x), op_safe_sqrt(v.y), op_safe_sqrt(v.z)); }
static v3 op_pow3(v3 v, float e) {
    return mk3(powf(op_max(v.x, 0.0f), e), powf(op_max(v.y, 0.0f), e), powf(op_max(v.z, 0.0f), e));
}

/* ---- transport context: the path throughput picks the wavelength channel */
static v3 op_channel_probs(v3 thr) {
    const v3 t = op_clamp_nonneg(thr);
    const float total = t.x + t.y + t.z;
    return total > 1e-8f ? op_div(t, total) : op_vec(1.0f / 3.0f);
}
static int op_sample_channel(v3 thr, float u) {
    const v3 p = op_channel_probs(thr);
    const float c = op_clamp(u, 0.0f, 0.99999994f);
    return c < p.x ? 0 : (c < p.x + p.y ? 1 : 2);
}

static v3 op_surface_transmission_color(const rto_openpbr* m) {
    if (m->geometry_thin_walled != 0 || m->transmission_depth <= 0.0f)
        return op_clamp_unit(ld3(m->transmission_color));
    return op_vec(1.0f);
}

/* ---- subsurface random walk (the medium the path carries inside) */
static float op_ss_albedo(float observed, float anisotropy) {
    const float color = op_sat(observed);
    const float g = op_clamp(anisotropy, -0.999f, 0.999f);
    const float s = op_max(0.0f, 4.09712f + 4.20863f * color
                                     - sqrtf(op_max(0.0f, 9.59217f + 41.6808f * color
                                                              + 17.7126f * color * color)));
    const float s2 = s * s;
    return op_sat((1.0f - s2) / op_max(1.0f - g * s2, 1e-6f));
}
static int op_ss_finite_mfp(const rto_openpbr* m) {
    const float radius = op_max(0.0f, m->subsurface_radius);
    const v3 sc = op_clamp_nonneg(ld3(m->subsurface_radius_scale));
    return radius * op_max(sc.x, op_max(sc.y, sc.z)) > 1e-6f;
}
static int op_ss_random_walk(const rto_openpbr* m) {
    return m->geometry_thin_walled == 0 && op_sat(m->subsurface_weight) > 0.0f
           && op_sat(m->base_metalness) < 1.0f && op_sat(m->transmission_weight) < 1.0f
           && op_ss_finite_mfp(m);
}
static int op_ss_thin_walled(const rto_openpbr* m) {
    return m->geometry_thin_walled != 0 && op_sat(m->subsurface_weight) > 0.0f;
}
static float op_ss_thin_prob(const rto_openpbr* m, int transmission) {
    const float g = op_clamp(m->subsurface_scatter_anisotropy, -1.0f, 1.0f);
    return 0.5f * (1.0f + (transmission ? g : -g));
}

typedef struct { v3 extinction, albedo; float g; int active; } op_medium;

static op_medium op_ss_medium(const rto_openpbr* m) {
    op_medium md;
    md.extinction = op_vec(0.0f);
    md.albedo = op_vec(0.0f);
    md.g = 0.0f;
    md.active = 0;
    if (!op_ss_random_walk(m)) return md;
    const float radius = op_max(0.0f, m->subsurface_radius);
    const v3 sc = op_clamp_nonneg(ld3(m->subsurface_radius_scale));
    const v3 mfp = mk3(op_max(radius * sc.x, 1e-6f), op_max(radius * sc.y, 1e-6f),
                       op_max(radius * sc.z, 1e-6f));
    md.extinction = op_recip(mfp);
    const v3 color = op_clamp_unit(ld3(m->subsurface_color));
    md.albedo = mk3(op_ss_albedo(color.x, m->subsurface_scatter_anisotropy),
                    op_ss_albedo(color.y, m->subsurface_scatter_anisotropy),
                    op_ss_albedo(color.z, m->subsurface_scatter_anisotropy));
    md.g = op_clamp(m->subsurface_scatter_anisotropy, -0.999f, 0.999f);
    md.active = 1;
    return md;
}
static v3 op_ss_channel_probs(const op_medium* md, v3 thr) {
    const v3 c = mul3(op_clamp_nonneg(thr), md->albedo);
    const float total = c.x + c.y + c.z;
    return total > 1e-12f ? op_div(c, total) : op_vec(1.0f / 3.0f);
}
static v3 op_ss_transmittance(const op_medium* md, float d) {
    const float sd = op_max(0.0f, d);
    return mk3(expf(-md->extinction.x * sd), expf(-md->extinction.y * sd),
               expf(-md->extinction.z * sd));
}
/* out: weight[3], distance, scattered */
static void op_ss_segment(const op_medium* md, float boundary, float u, v3 thr, float out[5]) {
    v3 weight = op_vec(1.0f);
    float dist = op_max(0.0f, boundary);
    int scattered = 0;
    if (md->active != 0 && dist > 0.0f) {
        const v3 p = op_ss_channel_probs(md, thr);
        const float choice = op_clamp(u, 0.0f, 0.99999994f);
        int channel = 2;
        float cu = 0.0f, cdf = 0.0f;
        for (int i = 0; i < 3; ++i) {
            const float pi = op_comp(p, i);
            if (choice < cdf + pi || i == 2) {
                channel = i;
                cu = (choice - cdf) / op_max(pi, 1e-8f);
                break;
            }
            cdf += pi;
        }
        const float ext = op_comp(md->extinction, channel);
        const float sampled = -logf(op_max(1.0f - op_sat(cu), 1e-7f)) / op_max(ext, 1e-6f);
        scattered = sampled < dist ? 1 : 0;
        dist = scattered ? sampled : dist;
        const v3 tr = op_ss_transmittance(md, dist);
        if (scattered) {
            const v3 density = mul3(md->extinction, tr);
            const float mix = dot3(p, density);
            weight = op_div(mul3(density, md->albedo), op_max(mix, 1e-12f));
        } else {
            const float surv = dot3(p, tr);
            weight = op_div(tr, op_max(surv, 1e-12f));
        }
    }
    out[0] = weight.x; out[1] = weight.y; out[2] = weight.z;
    out[3] = dist;
    out[4] = (float)scattered;
}

/* ---- frames */
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

static v3 op_sample_hg(v3 incident, float anisotropy, float u1, float u2) {
    const float g = op_clamp(anisotropy, -0.999f, 0.999f);
    float c = 1.0f - 2.0f * op_sat(u1);
    if (fabsf(g) > 1e-3f) {
        const float ratio = (1.0f - g * g) / op_max(1.0f - g + 2.0f * g * op_sat(u1), 1e-6f);
        c = op_clamp((1.0f + g * g - ratio * ratio) / (2.0f * g), -1.0f, 1.0f);
    }
    const float s = op_safe_sqrt(1.0f - c * c);
    float sp, cp;
    om_sincos_turn(op_sat(u2), &sp, &cp);
    const v3 local = mk3(s * cp, s * sp, c);
    const op_frame f = op_make_frame(op_normalize(incident));
    return op_normalize(op_to_world(&f, local));
}

/* ---- Fresnel */
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
static op_pol op_fresnel_polarized(float cosine, float eta) {
    const float c2 = op_sq(op_sat(cosine));
    const float s2 = 1.0f - c2;
    const float t0 = op_max(eta * eta - s2, 0.0f);
    const float t1 = t0 + c2;
    const float t2 = 2.0f * sqrtf(t0) * op_sat(cosine);
    const float rs = (t1 - t2) / op_max(t1 + t2, 1e-8f);
    const float t3 = c2 * t0 + s2 * s2;
    const float t4 = t2 * s2;
    const float rp = rs * (t3 - t4) / op_max(t3 + t4, 1e-8f);
    op_pol r = {rp, rs};
    return r;
}
static op_pol op_fresnel_real_polarized(float cosine, float rel) {
    const float c2 = op_sq(op_sat(cosine));
    const float s2 = 1.0f - c2;
    const float t0 = rel * rel - s2;
    const float mag = fabsf(t0);
    const float t1 = mag + c2;
    const float a = op_safe_sqrt(0.5f * (mag + t0));
    const float t2 = 2.0f * a * op_sat(cosine);
    const float rs = (t1 - t2) / op_max(t1 + t2, 1e-8f);
    const float t3 = c2 * mag + s2 * s2;
    const float t4 = t2 * s2;
    const float rp = rs * (t3 - t4) / op_max(t3 + t4, 1e-8f);
    op_pol r = {rp, rs};
    return r;
}
static op_pol op_real_phase(float cosine, float inc, float trans) {
    const float s2 = 1.0f - op_sq(op_sat(cosine));
    const float t2 = trans * trans, i2 = inc * inc;
    const float a = t2 - i2 * s2;
    const float u = op_safe_sqrt(0.5f * (a + fabsf(a)));
    const float sd = u * u - op_sq(inc * cosine);
    const float pd = op_sq(t2 * cosine) - i2 * u * u;
    op_pol r = {pd < 0.0f ? OP_PI : 0.0f, sd < 0.0f ? OP_PI : 0.0f};
    return r;
}

/* ---- thin film: Airy summation with the Belcour-Barla XYZ sensitivity fit */
static v3 op_film_sensitivity(float opd, v3 shift) {
    const float phase = 2.0f * OP_PI * opd;
    const v3 val = mk3(5.4856e-13f, 4.4201e-13f, 5.2481e-13f);
    const v3 pos = mk3(1.6810e6f, 1.7953e6f, 2.2084e6f);
    const v3 var = mk3(4.3278e9f, 9.3046e9f, 6.6121e9f);
    v3 xyz = mk3(val.x * sqrtf(2.0f * OP_PI * var.x) * cosf(pos.x * phase + shift.x) * expf(-var.x * phase * phase),
                 val.y * sqrtf(2.0f * OP_PI * var.y) * cosf(pos.y * phase + shift.y) * expf(-var.y * phase * phase),
                 val.z * sqrtf(2.0f * OP_PI * var.z) * cosf(pos.z * phase + shift.z) * expf(-var.z * phase * phase));
    xyz.x += 9.7470e-14f * sqrtf(2.0f * OP_PI * 4.5282e9f) * cosf(2.2399e6f * phase + shift.x)
             * expf(-4.5282e9f * phase * phase);
    return op_div(xyz, 1.0685e-7f);
}
static v3 op_xyz_to_rgb(v3 c) {
    return mk3(2.3706743f * c.x - 0.9000405f * c.y - 0.4706338f * c.z,
               -0.5138850f * c.x + 1.4253036f * c.y + 0.0885814f * c.z,
               0.0052982f * c.x - 0.0146949f * c.y + 1.0093968f * c.z);
}
static float op_f0_to_ior(float f0) {
    const float r = sqrtf(op_clamp(f0, 0.01f, 0.99f));
    return (1.0f + r) / op_max(1.0f - r, 1e-5f);
}
static float op_film_cos_t(float cosine, float inc, float film) {
    const float ratio = inc / op_max(film, 1e-4f);
    const float s2 = (1.0f - op_sq(op_sat(cosine))) * ratio * ratio;
    return s2 < 1.0f ? op_safe_sqrt(1.0f - s2) : 0.0f;
}
static v3 op_film_airy(float cosine, float thick_nm, float inc_ior, float film_ior, v3 sub_ior,
    v3 gen_refl, int gen_schlick) {
    const float eta1 = op_max(inc_ior, 1e-4f);
    const float eta2 = op_max(film_ior, 1e-4f);
    const float ct = op_sat(fabsf(cosine));
    const float ctt = op_film_cos_t(ct, eta1, eta2);
    if (ctt <= 0.0f) return op_vec(1.0f);
    const op_pol r12 = op_fresnel_polarized(ct, eta2 / eta1);
    const op_pol t121 = {1.0f - r12.p, 1.0f - r12.s};
    v3 r23p = op_vec(0.0f), r23s = op_vec(0.0f), ph23p = op_vec(0.0f), ph23s = op_vec(0.0f);
    for (int c = 0; c < 3; ++c) {
        const float eta3 = op_comp(sub_ior, c);
        if (gen_schlick) {
            const float refl = 0.5f * op_sat(op_comp(gen_refl, c));
            const float ph = eta3 < eta2 ? OP_PI : 0.0f;
            op_set_comp(&r23p, c, refl);
            op_set_comp(&r23s, c, refl);
            op_set_comp(&ph23p, c, ph);
            op_set_comp(&ph23s, c, ph);
        } else {
            const op_pol refl = op_fresnel_real_polarized(ctt, eta3 / eta2);
            const op_pol ph = op_real_phase(ctt, eta2, eta3);
            op_set_comp(&r23p, c, refl.p);
            op_set_comp(&r23s, c, refl.s);
            op_set_comp(&ph23p, c, ph.p);
            op_set_comp(&ph23s, c, ph.s);
        }
    }
    const float brewster = eta1 / sqrtf(eta1 * eta1 + eta2 * eta2);
    const op_pol ph21 = {ct < brewster ? 0.0f : OP_PI, OP_PI};
    const v3 r123p = mk3(sqrtf(op_max(r12.p * r23p.x, 0.0f)), sqrtf(op_max(r12.p * r23p.y, 0.0f)),
                         sqrtf(op_max(r12.p * r23p.z, 0.0f)));
    const v3 r123s = mk3(sqrtf(op_max(r12.s * r23s.x, 0.0f)), sqrtf(op_max(r12.s * r23s.y, 0.0f)),
                         sqrtf(op_max(r12.s * r23s.z, 0.0f)));
    const float opd = 2.0f * eta2 * ctt * thick_nm * 1e-9f;
    /* p polarization (the reference names the transmitted sum rs) */
    const v3 sp = mk3(t121.p * t121.p * r23p.x / op_max(1.0f - r12.p * r23p.x, 1e-6f),
                      t121.p * t121.p * r23p.y / op_max(1.0f - r12.p * r23p.y, 1e-6f),
                      t121.p * t121.p * r23p.z / op_max(1.0f - r12.p * r23p.z, 1e-6f));
    v3 result = add3(op_vec(r12.p), sp);
    v3 coeff = sub3(sp, op_vec(t121.p));
    for (int order = 1; order <= 2; ++order) {
        coeff = mul3(coeff, r123p);
        const v3 shift = scl3(add3(ph23p, op_vec(ph21.p)), (float)order);
        result = add3(result, mul3(coeff, scl3(op_film_sensitivity((float)order * opd, shift), 2.0f)));
    }
    const v3 ss = mk3(t121.s * t121.s * r23s.x / op_max(1.0f - r12.s * r23s.x, 1e-6f),
                      t121.s * t121.s * r23s.y / op_max(1.0f - r12.s * r23s.y, 1e-6f),
                      t121.s * t121.s * r23s.z / op_max(1.0f - r12.s * r23s.z, 1e-6f));
    result = add3(result, add3(op_vec(r12.s), ss));
    coeff = sub3(ss, op_vec(t121.s));
    for (int order = 1; order <= 2; ++order) {
        coeff = mul3(coeff, r123s);
        const v3 shift = scl3(add3(ph23s, op_vec(ph21.s)), (float)order);
        result = add3(result, mul3(coeff, scl3(op_film_sensitivity((float)order * opd, shift), 2.0f)));
    }
    return op_clamp_unit(op_xyz_to_rgb(scl3(result, 0.5f)));
}

/* ---- IOR and dispersion */
static float op_modulated_ior(const rto_openpbr* m) {
    const float eta = op_max(m->specular_ior, 1e-4f);
    const float ratio = (eta - 1.0f) / (eta + 1.0f);
    const float f0 = op_clamp(op_max(0.0f, m->specular_weight) * ratio * ratio, 0.0f, 0.9999f);
    const float sign = eta < 1.0f ? -1.0f : (eta > 1.0f ? 1.0f : 0.0f);
    const float eps = sign * sqrtf(f0);
    return op_max((1.0f + eps) / op_max(1.0f - eps, 1e-4f), 1e-4f);
}
static int op_dispersion(const rto_openpbr* m) {
    return m->transmission_dispersion_scale > 0.0f && op_sat(m->base_metalness) < 1.0f;
}
static float op_dispersion_ior1(float ior0, float scale_v, float abbe, float wl) {
    const float scale = op_sat(scale_v);
    if (scale <= 0.0f || fabsf(ior0 - 1.0f) <= 1e-8f) return ior0;
    const int recip = ior0 < 1.0f;
    const float ior = recip ? 1.0f / op_max(ior0, 1e-6f) : ior0;
    const float vd = op_max(abbe, 1e-4f) / scale;
    const float inv_f2 = 1.0f / op_sq(OP_FRAUNHOFER_F);
    const float inv_c2 = 1.0f / op_sq(OP_FRAUNHOFER_C);
    const float b = (ior - 1.0f) / (vd * (inv_f2 - inv_c2));
    const float a = ior - b / op_sq(OP_FRAUNHOFER_D);
    const float adj = a + b / op_sq(op_max(wl, 1e-4f));
    return recip ? 1.0f / op_max(adj, 1e-6f) : adj;
}
static v3 op_dispersion_ior(const rto_openpbr* m) {
    const float ior = op_modulated_ior(m);
    if (!op_dispersion(m)) return op_vec(ior);
    const float s = m->transmission_dispersion_scale, ab = m->transmission_dispersion_abbe_number;
    return mk3(op_dispersion_ior1(ior, s, ab, OP_FRAUNHOFER_C),
               op_dispersion_ior1(ior, s, ab, OP_FRAUNHOFER_D),
               op_dispersion_ior1(ior, s, ab, OP_FRAUNHOFER_F));
}

/* ---- roughness and GGX */
static op_rough op_roughness2(float rv, float av) {
    const float r = op_sat(rv);
    const float aniso = op_sat(av);
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
static op_rough op_spec_rough(const rto_openpbr* m) {
    return op_roughness2(op_eff_spec_roughness(m), m->specular_roughness_anisotropy);
}
static op_rough op_coat_rough(const rto_openpbr* m) {
    return op_roughness2(m->coat_roughness, m->coat_roughness_anisotropy);
}
static int op_smooth(op_rough r) { return op_max(r.ax, r.ay) < 1e-3f; }
static int op_smooth_m(const rto_openpbr* m) { return op_smooth(op_spec_rough(m)); }
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
/* refraction of wo about wm with relative IOR eta; flips wm (and eta) for a
 * direction below it; returns 0 at total internal reflection */
static int op_refract(v3 wo, v3 wm, float eta, float* eta_path, v3* wi) {
    float ci = dot3(wm, wo);
    *eta_path = op_max(eta, 1e-4f);
    if (ci < 0.0f) { wm = neg3(wm); ci = -ci; *eta_path = 1.0f / *eta_path; }
    const float s2i = op_max(0.0f, 1.0f - ci * ci);
    const float s2t = s2i / (*eta_path * *eta_path);
    if (s2t >= 1.0f) return 0;
    const float ct = op_safe_sqrt(1.0f - s2t);
    *wi = op_normalize(add3(scl3(wo, -1.0f / *eta_path), scl3(wm, ci / *eta_path - ct)));
    return 1;
}

/* ---- coat */
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
    return op_lerp(op_vec(1.0f), op_pow3(op_sqrt3(tint), ds), presence);
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

/* ---- fuzz: LTC sheen */
static float op_fuzz_albedo(float cosine, float rv) {
    const float x = op_sat(cosine);
    const float y = op_clamp(rv, 0.01f, 1.0f);
    const float s = y * (0.0206607f + 1.58491f * y) / (0.0379424f + y * (1.32227f + y));
    const float y2 = y * y;
    const float mm = y * (-0.193854f + y * (-1.14885f + y * (1.7932f - 0.95943f * y2))) / (0.046391f + y);
    const float o = y * (0.000654023f + (-0.0207818f + 0.119681f * y) * y) / (1.26264f + y * (-1.92021f + y));
    const float nrm = (x - mm) / op_max(s, 1e-5f);
    const float gauss = expf(-0.5f * nrm * nrm) / (op_max(s, 1e-5f) * sqrtf(2.0f * OP_PI));
    return op_sat(gauss + o);
}
static op_frame op_fuzz_frame(v3 wo) {
    const v3 proj = mk3(wo.x, wo.y, 0.0f);
    const v3 t = lensq3(proj) > 1e-12f ? op_normalize(proj) : mk3(1.0f, 0.0f, 0.0f);
    op_frame f;
    f.t = t;
    f.b = mk3(-t.y, t.x, 0.0f);
    f.n = mk3(0.0f, 0.0f, 1.0f);
    return f;
}
static v3 op_fuzz_ltc(float cosine, float rv) {
    const float x = op_sat(cosine);
    const float y = op_clamp(rv, 0.01f, 1.0f);
    const float ainv = (2.58126f * x + 0.813703f * y) * y / (1.0f + 0.310327f * x * x + 2.60994f * x * y);
    const float binv = op_safe_sqrt(1.0f - x) * (y - 1.0f) * y * y * y
                       / (0.0000254053f + 1.71228f * x - 1.71506f * x * y + 1.34174f * y * y);
    return mk3(op_max(ainv, 1e-5f), binv, op_fuzz_albedo(x, y));
}
static float op_fuzz_pdf(const rto_openpbr* m, v3 wo, v3 wi) {
    if (wo.z <= 0.0f || wi.z <= 0.0f) return 0.0f;
    const op_frame f = op_fuzz_frame(wo);
    const v3 al = op_to_local(&f, wi);
    const v3 c = op_fuzz_ltc(wo.z, m->fuzz_roughness);
    const v3 tr = mk3(c.x * al.x + c.y * al.z, c.x * al.y, al.z);
    const float l2 = lensq3(tr);
    if (l2 <= 1e-16f || tr.z <= 0.0f) return 0.0f;
    const float js = c.x / l2;
    return tr.z * OP_INV_PI * js * js;
}
static v3 op_sample_fuzz(const rto_openpbr* m, v3 wo, float u1, float u2) {
    const float radius = sqrtf(op_sat(u1));
    float sp, cp;
    om_sincos_turn(op_sat(u2), &sp, &cp);
    const v3 cs = mk3(radius * cp, radius * sp, op_safe_sqrt(1.0f - radius * radius));
    const v3 c = op_fuzz_ltc(wo.z, m->fuzz_roughness);
    const float ia = 1.0f / c.x;
    const v3 al = op_normalize(mk3(cs.x * ia - cs.z * c.y * ia, cs.y * ia, cs.z));
    const op_frame f = op_fuzz_frame(wo);
    return op_normalize(op_to_world(&f, al));
}
static v3 op_fuzz_value(const rto_openpbr* m, v3 wo, v3 wi) {
    const float presence = op_sat(m->fuzz_weight);
    if (presence <= 0.0f || wo.z <= 0.0f || wi.z <= 0.0f) return op_vec(0.0f);
    const float refl = op_fuzz_albedo(wo.z, m->fuzz_roughness);
    const float dens = op_fuzz_pdf(m, wo, wi);
    return scl3(op_clamp_unit(ld3(m->fuzz_color)), presence * refl * dens / op_max(wi.z, 1e-6f));
}
static float op_fuzz_base_scale(const rto_openpbr* m, float ocos, float icos, int exterior) {
    if (!exterior) return 1.0f;
    const float presence = op_sat(m->fuzz_weight);
    const float out = 1.0f - presence * op_fuzz_albedo(ocos, m->fuzz_roughness);
    const float in = icos > 0.0f ? 1.0f - presence * op_fuzz_albedo(icos, m->fuzz_roughness) : 1.0f;
    return op_sat(out) * op_sat(in);
}

/* ---- layering */
static v3 op_coat_base_scale(const rto_openpbr* m, float ocos, float icos, int exterior) {
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
static v3 op_substrate_scale(const rto_openpbr* m, float ocos, float icos, int exterior) {
    return scl3(op_coat_base_scale(m, ocos, icos, exterior), op_fuzz_base_scale(m, ocos, icos, exterior));
}
static v3 op_coat_value(const rto_openpbr* m, v3 wo, v3 wi) {
    const float presence = op_sat(m->coat_weight);
    if (presence <= 0.0f || wo.z <= 0.0f || wi.z <= 0.0f) return op_vec(0.0f);
    const op_rough r = op_coat_rough(m);
    if (op_smooth(r)) return op_vec(0.0f);
    const v3 sm = add3(wi, wo);
    if (lensq3(sm) <= 1e-16f) return op_vec(0.0f);
    v3 wm = op_normalize(sm);
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
    const v3 sm = add3(wi, wo);
    if (lensq3(sm) <= 1e-16f) return 0.0f;
    v3 wm = op_normalize(sm);
    if (wm.z < 0.0f) wm = neg3(wm);
    const float wodm = dot3(wo, wm);
    if (wodm <= 0.0f) return 0.0f;
    return op_vndf_pdf(r, wo, wm) / op_max(4.0f * fabsf(wodm), 1e-8f);
}

/* ---- base Fresnel terms */
static float op_dielectric_fresnel(const rto_openpbr* m, float cosine, int entering) {
    float eta = op_modulated_ior(m);
    if (m->geometry_thin_walled == 0 && !entering) eta = 1.0f / eta;
    float f = op_fresnel_dielectric(cosine, eta);
    if (m->geometry_thin_walled != 0 && f < 1.0f) {
        const float t = 1.0f - f;
        f += t * t * f / op_max(1.0f - f * f, 1e-6f);
    }
    return op_sat(f);
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
static float op_weighted_coat_ior(const rto_openpbr* m) {
    return 1.0f + op_sat(m->coat_weight) * (op_max(m->coat_ior, 1e-4f) - 1.0f);
}
static float op_window1(float r) {
    return op_sat(r + op_sq(1.0f - r) * r / op_max(1.0f - r * r, 1e-6f));
}
static v3 op_window(v3 refl) {
    const v3 r = op_clamp_unit(refl);
    return mk3(op_window1(r.x), op_window1(r.y), op_window1(r.z));
}
static v3 op_apply_film(const rto_openpbr* m, float cosine, v3 base, float inc_ior, v3 sub_ior,
    v3 gen_refl, int gen_schlick) {
    const float w = op_sat(m->thin_film_weight);
    const float nm = op_max(m->thin_film_thickness, 0.0f) * 1000.0f;
    if (w <= 0.0f || nm <= 0.0f) return base;
    v3 film = op_film_airy(cosine, nm, inc_ior, m->thin_film_ior, sub_ior, gen_refl, gen_schlick);
    if (m->geometry_thin_walled != 0 && !gen_schlick) film = op_window(film);
    return op_lerp(base, film, w);
}
static v3 op_dielectric_reflectance(const rto_openpbr* m, float cosine, int entering) {
    if (!op_dispersion(m)) {
        const float base = op_dielectric_fresnel(m, cosine, entering);
        const float bior = op_modulated_ior(m);
        const float cior = op_weighted_coat_ior(m);
        const float inc = entering ? cior : bior;
        const float sub = entering ? bior : cior;
        return op_apply_film(m, cosine, op_vec(base), inc, op_vec(sub), op_vec(0.0f), 0);
    }
    const v3 iors = op_dispersion_ior(m);
    const v3 eta = entering ? iors : op_recip(iors);
    v3 base = mk3(op_fresnel_dielectric(cosine, eta.x), op_fresnel_dielectric(cosine, eta.y),
                  op_fresnel_dielectric(cosine, eta.z));
    if (m->geometry_thin_walled != 0) base = op_window(base);
    const float cior = op_weighted_coat_ior(m);
    if (entering) return op_apply_film(m, cosine, base, cior, iors, op_vec(0.0f), 0);
    v3 r = op_vec(0.0f);
    for (int c = 0; c < 3; ++c) {
        const v3 film = op_apply_film(m, cosine, base, op_comp(iors, c), op_vec(cior), op_vec(0.0f), 0);
        op_set_comp(&r, c, op_comp(film, c));
    }
    return r;
}
static v3 op_metal_reflectance(const rto_openpbr* m, float cosine) {
    const v3 base = op_metal_fresnel(m, cosine);
    const float inc = op_weighted_coat_ior(m);
    const float ctt = op_film_cos_t(cosine, inc, m->thin_film_ior);
    const v3 at_film = op_metal_fresnel(m, ctt);
    const v3 normal = op_metal_fresnel(m, 1.0f);
    const v3 sub = mk3(inc * op_f0_to_ior(normal.x), inc * op_f0_to_ior(normal.y),
                       inc * op_f0_to_ior(normal.z));
    return op_apply_film(m, cosine, base, inc, sub, at_film, 1);
}
static v3 op_dielectric_albedo_rgb(const rto_openpbr* m, float cosine, int entering) {
    if ((m->thin_film_weight <= 0.0f || m->thin_film_thickness <= 0.0f) && !op_dispersion(m)) {
        const op_rough r = op_spec_rough(m);
        const float alpha = op_avg_alpha(r);
        if (op_smooth_m(m)) return op_vec(op_dielectric_fresnel(m, cosine, entering));
        const float c0 = op_dielectric_fresnel(m, 1.0f, entering);
        if (c0 <= 1e-8f) return op_vec(0.0f);
        const float ss = op_ggx_albedo(cosine, alpha, c0, 1.0f);
        const float white = op_ggx_albedo(cosine, alpha, 1.0f, 1.0f);
        const float df = op_dielectric_fresnel(m, cosine, entering);
        const float comp = 1.0f + df * (1.0f - white) / op_max(white, 1e-4f);
        return op_vec(op_sat(ss * comp));
    }
    const op_rough r = op_spec_rough(m);
    const float alpha = op_avg_alpha(r);
    const v3 mirror = op_dielectric_reflectance(m, cosine, entering);
    if (op_smooth_m(m)) return mirror;
    const v3 c0 = op_dielectric_reflectance(m, 1.0f, entering);
    const v3 rough = mk3(op_ggx_albedo(cosine, alpha, c0.x, 1.0f), op_ggx_albedo(cosine, alpha, c0.y, 1.0f),
                         op_ggx_albedo(cosine, alpha, c0.z, 1.0f));
    return op_lerp(mirror, rough, sqrtf(op_sat(alpha)));
}
static v3 op_reflection_fresnel(const rto_openpbr* m, float mcos, float ocos, int entering) {
    const float metal = op_sat(m->base_metalness);
    const float alpha = op_avg_alpha(op_spec_rough(m));
    const v3 df = op_dielectric_reflectance(m, mcos, entering);
    const v3 dsingle = mul3(entering || m->geometry_thin_walled != 0 ? op_clamp_unit(ld3(m->specular_color))
                                                                     : op_vec(1.0f), df);
    const v3 dielectric = mul3(dsingle, op_energy_comp(ocos, alpha, df));
    const v3 msingle = op_metal_reflectance(m, mcos);
    const v3 metal_v = mul3(msingle, op_energy_comp(ocos, alpha, msingle));
    return op_lerp(dielectric, metal_v, metal);
}

/* ---- lobe selection */
static op_probs op_lobe_probs(const rto_openpbr* m, float ocos, int exterior) {
    op_probs p = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    const float metal = op_sat(m->base_metalness);
    const float trans = op_sat(m->transmission_weight);
    const float bw = op_sat(m->base_weight);
    const float msw = op_sat(m->specular_weight);
    const float eta = op_modulated_ior(m);
    const float er = (eta - 1.0f) / (eta + 1.0f);
    const float df0 = er * er;
    float dsr = df0;
    float msr = msw * op_lum(scl3(op_clamp_unit(ld3(m->base_color)), bw));
    if ((m->thin_film_weight > 0.0f && m->thin_film_thickness > 0.0f) || op_dispersion(m)) {
        dsr = op_lum(op_dielectric_reflectance(m, ocos, 1));
        msr = op_lum(op_metal_reflectance(m, ocos));
    }
    const float opaque = (1.0f - metal) * (1.0f - trans) * op_sat(1.0f - dsr);
    const float ssw = op_sat(m->subsurface_weight);
    float diffuse = 0.0f, sss = 0.0f;
    if (ssw <= 0.0f) {
        diffuse = opaque * bw * op_lum(op_clamp_unit(ld3(m->base_color)));
    } else {
        const int rw = op_ss_random_walk(m);
        const int tw = op_ss_thin_walled(m);
        const float local_ss = rw || tw ? 0.0f : ssw;
        const v3 ldc = add3(scl3(op_clamp_unit(ld3(m->base_color)), bw * (1.0f - ssw)),
                            scl3(op_clamp_unit(ld3(m->subsurface_color)), local_ss));
        diffuse = opaque * op_lum(ldc);
        if (rw || tw) sss = opaque * ssw * op_lum(op_clamp_unit(ld3(m->subsurface_color)));
    }
    const float drs = (1.0f - metal) * dsr * op_lum(op_clamp_unit(ld3(m->specular_color)));
    const float mrs = metal * msr;
    const float ts = (1.0f - metal) * trans * op_sat(1.0f - dsr) * op_lum(op_surface_transmission_color(m));
    const float coat = exterior ? op_sat(m->coat_weight) * op_coat_albedo(m, ocos) : 0.0f;
    const float fuzz = exterior ? op_sat(m->fuzz_weight) * op_fuzz_albedo(ocos, m->fuzz_roughness)
                                      * op_lum(op_clamp_unit(ld3(m->fuzz_color)))
                                : 0.0f;
    const float scale = op_sat(1.0f - coat) * op_sat(1.0f - fuzz);
    const float sd = diffuse * scale;
    const float sssd = sss * scale;
    const float rs = (drs + mrs) * scale;
    const float st = ts * scale;
    const float total = sd + sssd + rs + st + coat + fuzz;
    if (total <= 1e-10f) return p;
    p.diffuse = sd / total;
    p.subsurface = sssd / total;
    p.reflection = rs / total;
    p.transmission = st / total;
    p.coat = coat / total;
    p.fuzz = fuzz / total;
    return p;
}

/* ---- diffuse */
static float op_on_dir_albedo(float cosine, float rv) {
    const float k1 = 0.5f - 2.0f / (3.0f * OP_PI);
    const float c = op_clamp(cosine, 1e-6f, 1.0f);
    const float sigma = op_sat(rv);
    const float a = 1.0f / (1.0f + k1 * sigma);
    const float b = sigma * a;
    const float s = op_safe_sqrt(1.0f - c * c);
    const float g = s * (acosf(c) - s * c) + 2.0f * ((s / c) * (1.0f - s * s * s) - s) / 3.0f;
    return op_sat(a + b * g * OP_INV_PI);
}
static float op_on_avg_albedo(float rv) {
    const float k1 = 0.5f - 2.0f / (3.0f * OP_PI);
    const float k2 = 2.0f / 3.0f - 28.0f / (15.0f * OP_PI);
    const float sigma = op_sat(rv);
    const float a = 1.0f / (1.0f + k1 * sigma);
    return op_sat(a * (1.0f + k2 * sigma));
}
static v3 op_diffuse(const rto_openpbr* m, v3 wo, v3 wi) {
    if (wo.z <= 0.0f || wi.z <= 0.0f) return op_vec(0.0f);
    const float metal = op_sat(m->base_metalness);
    const float trans = op_sat(m->transmission_weight);
    const float ssw = op_sat(m->subsurface_weight);
    float mix = op_sat(m->base_weight);
    v3 color = op_clamp_unit(ld3(m->base_color));
    if (ssw > 0.0f) {
        const float bmix = op_sat(m->base_weight) * (1.0f - ssw);
        const float smix = op_ss_random_walk(m) || op_ss_thin_walled(m) ? 0.0f : ssw;
        mix = bmix + smix;
        color = op_clamp_unit(op_div(add3(scl3(op_clamp_unit(ld3(m->base_color)), bmix),
                                          scl3(op_clamp_unit(ld3(m->subsurface_color)), smix)),
                                     op_max(mix, 1e-8f)));
    }
    const float coeff = mix * (1.0f - metal) * (1.0f - trans);
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
static v3 op_thin_ss_value(const rto_openpbr* m, v3 wo, v3 wi_same, int transmission) {
    if (!op_ss_thin_walled(m)) return op_vec(0.0f);
    rto_openpbr d = *m;
    d.base_weight = op_sat(m->subsurface_weight) * op_ss_thin_prob(m, transmission);
    const v3 sc = op_clamp_unit(ld3(m->subsurface_color));
    d.base_color[0] = sc.x; d.base_color[1] = sc.y; d.base_color[2] = sc.z;
    d.subsurface_weight = 0.0f;
    return op_diffuse(&d, wo, wi_same);
}

/* ---- evaluate */
static void op_eval(const rto_openpbr* m, v3 n, v3 wo_world, v3 wi_world, v3 thr, v3* value, float* pdf) {
    *value = op_vec(0.0f);
    *pdf = 0.0f;
    const v3 won = op_normalize(wo_world);
    const v3 win = op_normalize(wi_world);
    if (lensq3(won) <= 0.0f || lensq3(win) <= 0.0f) return;
    op_frame f = op_make_frame(n);
    const int entering = m->geometry_thin_walled != 0 || dot3(f.n, won) >= 0.0f;
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
        val = add3(op_diffuse(m, wo, wi), op_thin_ss_value(m, wo, wi, 0));
        val = mul3(val, op_clamp_unit(att));
        pd = p.diffuse * wi.z * OP_INV_PI;
        if (op_ss_thin_walled(m)) pd += p.subsurface * op_ss_thin_prob(m, 0) * wi.z * OP_INV_PI;
        const v3 sm = add3(wi, wo);
        if (!smooth && lensq3(sm) > 1e-16f) {
            v3 wm = op_normalize(sm);
            if (wm.z < 0.0f) wm = neg3(wm);
            const float wodm = dot3(wo, wm);
            if (wodm > 0.0f) {
                const float D = op_ggx_d(r, wm);
                const float G = op_ggx_g(r, wo, wi);
                const v3 F = op_reflection_fresnel(m, wodm, wo.z, entering);
                val = add3(val, scl3(F, D * G / op_max(4.0f * wo.z * wi.z, 1e-8f)));
                pd += p.reflection * op_vndf_pdf(r, wo, wm) / op_max(4.0f * fabsf(wodm), 1e-8f);
            }
        }
        val = mul3(val, op_substrate_scale(m, wo.z, wi.z, exterior));
        if (exterior) {
            const v3 coat = scl3(op_coat_value(m, wo, wi), op_fuzz_base_scale(m, wo.z, wi.z, 1));
            val = add3(val, coat);
            val = add3(val, op_fuzz_value(m, wo, wi));
            pd += p.coat * op_coat_pdf(m, wo, wi);
            pd += p.fuzz * op_fuzz_pdf(m, wo, wi);
        }
    } else if (op_ss_thin_walled(m) && p.subsurface > 0.0f) {
        const v3 wir = mk3(wi.x, wi.y, -wi.z);
        const v3 att = sub3(op_vec(1.0f), op_dielectric_albedo_rgb(m, wo.z, 1));
        val = op_thin_ss_value(m, wo, wir, 1);
        val = mul3(val, mul3(op_clamp_unit(att), op_substrate_scale(m, wo.z, wi.z, 1)));
        pd = p.subsurface * op_ss_thin_prob(m, 1) * fabsf(wi.z) * OP_INV_PI;
    } else if (!smooth && m->geometry_thin_walled == 0 && p.transmission > 0.0f) {
        const float scale = (1.0f - op_sat(m->base_metalness)) * op_sat(m->transmission_weight);
        if (!op_dispersion(m)) {
            const float eta = op_modulated_ior(m);
            const float ep = entering ? eta : 1.0f / eta;
            const v3 sm = add3(scl3(wi, ep), wo);
            if (lensq3(sm) > 1e-16f) {
                v3 wm = op_normalize(sm);
                if (wm.z < 0.0f) wm = neg3(wm);
                const float widm = dot3(wi, wm);
                const float wodm = dot3(wo, wm);
                if (widm * wi.z > 0.0f && wodm * wo.z > 0.0f) {
                    const v3 refl = op_dielectric_reflectance(m, wodm, entering);
                    const float hd = widm + wodm / ep;
                    const float d2 = hd * hd;
                    if (d2 > 1e-16f) {
                        const float D = op_ggx_d(r, wm);
                        const float G = op_ggx_g(r, wo, wi);
                        float t = D * G * fabsf(widm * wodm / (wi.z * wo.z * d2));
                        t /= ep * ep;
                        t *= scale;
                        val = scl3(mul3(op_surface_transmission_color(m),
                                        op_clamp_unit(sub3(op_vec(1.0f), refl))), t);
                        const float deriv = fabsf(widm) / d2;
                        pd = p.transmission * op_vndf_pdf(r, wo, wm) * deriv;
                    }
                }
            }
        } else {
            const v3 iors = op_dispersion_ior(m);
            const v3 eps = entering ? iors : op_recip(iors);
            const v3 tc = op_surface_transmission_color(m);
            v3 cond = op_vec(0.0f);
            for (int c = 0; c < 3; ++c) {
                const float ep = op_comp(eps, c);
                const v3 sm = add3(scl3(wi, ep), wo);
                if (lensq3(sm) <= 1e-16f) continue;
                v3 wm = op_normalize(sm);
                if (wm.z < 0.0f) wm = neg3(wm);
                const float widm = dot3(wi, wm);
                const float wodm = dot3(wo, wm);
                if (widm * wi.z <= 0.0f || wodm * wo.z <= 0.0f) continue;
                const float hd = widm + wodm / ep;
                const float d2 = hd * hd;
                if (d2 <= 1e-16f) continue;
                const float D = op_ggx_d(r, wm);
                const float G = op_ggx_g(r, wo, wi);
                float t = D * G * fabsf(widm * wodm / (wi.z * wo.z * d2));
                t *= scale / (ep * ep);
                const v3 refl = op_dielectric_reflectance(m, wodm, entering);
                op_set_comp(&val, c, op_comp(tc, c) * op_sat(1.0f - op_comp(refl, c)) * t);
                const float deriv = fabsf(widm) / d2;
                op_set_comp(&cond, c, op_vndf_pdf(r, wo, wm) * deriv);
            }
            pd = p.transmission * dot3(cond, op_channel_probs(thr));
        }
        val = mul3(val, op_substrate_scale(m, wo.z, wi.z, exterior));
    }
    *value = scl3(val, opacity);
    *pdf = pd * opacity;
}

v3 om_openpbr_emission(const rto_openpbr* m) {
    v3 r = scl3(op_clamp_nonneg(ld3(m->emission_color)),
                op_max(0.0f, m->emission_luminance) * op_sat(m->geometry_opacity));
    const float ct = 1.0f - op_sat(m->coat_weight) * op_coat_albedo(m, 1.0f);
    r = mul3(r, scl3(op_coat_passage(m, 1.0f), op_sat(ct)));
    const float ft = 1.0f - op_sat(m->fuzz_weight) * op_fuzz_albedo(1.0f, m->fuzz_roughness);
    return scl3(r, op_sat(ft));
}

/* ---- sample */
typedef struct { v3 wi, value, weight; float pdf, discrete_pdf; int event, delta, valid; } op_sample;

static op_sample op_sample_none(void) {
    op_sample s;
    s.wi = s.value = s.weight = op_vec(0.0f);
    s.pdf = s.discrete_pdf = 0.0f;
    s.event = s.delta = s.valid = 0;
    return s;
}
static op_sample op_finish(const rto_openpbr* m, v3 n, const op_frame* f, v3 won, v3 wi, int event, v3 thr) {
    op_sample s = op_sample_none();
    s.wi = op_normalize(op_to_world(f, wi));
    s.event = event;
    v3 val;
    float pdf;
    op_eval(m, n, won, s.wi, thr, &val, &pdf);
    if (pdf <= 1e-12f) return op_sample_none();
    s.value = val;
    s.pdf = pdf;
    s.weight = scl3(val, fabsf(dot3(f->n, s.wi)) / pdf);
    s.valid = 1;
    return s;
}
static op_sample op_delta(const op_frame* f, v3 wi, v3 coeff, float dp, int event, int normalize) {
    op_sample s = op_sample_none();
    s.wi = normalize ? op_normalize(op_to_world(f, wi)) : op_to_world(f, wi);
    s.discrete_pdf = dp;
    if (dp <= 0.0f) return op_sample_none();
    s.weight = op_div(coeff, dp);
    s.event = event;
    s.delta = 1;
    s.valid = 1;
    return s;
}
static v3 op_cos_hemisphere(float u1, float u2) {
    const float radius = sqrtf(op_sat(u1));
    float sp, cp;
    om_sincos_turn(op_sat(u2), &sp, &cp);
    return mk3(radius * cp, radius * sp, op_safe_sqrt(1.0f - radius * radius));
}
static rto_openpbr op_boundary_material(const rto_openpbr* m) {
    rto_openpbr b = *m;
    b.base_weight = 0.0f;
    b.base_metalness = 0.0f;
    b.transmission_weight = 1.0f;
    b.transmission_color[0] = b.transmission_color[1] = b.transmission_color[2] = 1.0f;
    b.transmission_depth = 0.0f;
    b.transmission_dispersion_scale = 0.0f;
    b.subsurface_weight = 0.0f;
    b.fuzz_weight = 0.0f;
    b.coat_weight = 0.0f;
    b.geometry_opacity = 1.0f;
    return b;
}
/* the boundary of the subsurface medium: glossy reflection or transmission
 * of a clear dielectric with the material's IOR and roughness */
static op_sample op_sample_boundary(const rto_openpbr* m, v3 n, v3 wo_world, float ul, float u1, float u2, v3 thr) {
    const v3 won = op_normalize(wo_world);
    if (lensq3(won) <= 0.0f) return op_sample_none();
    op_frame f = op_make_frame(n);
    const int entering = dot3(f.n, won) >= 0.0f;
    f = op_face(f, won);
    const v3 wo = op_to_local(&f, won);
    if (wo.z <= 1e-7f) return op_sample_none();
    const op_probs p = op_lobe_probs(m, wo.z, 0);
    const float choice = op_clamp(ul, 0.0f, 0.99999994f);
    v3 wi;
    int event;
    const int smooth = op_smooth_m(m);
    if (choice < p.reflection) {
        if (smooth) {
            wi = mk3(-wo.x, -wo.y, wo.z);
            return op_delta(&f, wi, op_reflection_fresnel(m, wo.z, wo.z, entering), p.reflection, 2, 1);
        }
        const v3 wm = op_sample_vndf(op_spec_rough(m), wo, u1, u2);
        wi = op_reflect(wo, wm);
        if (wi.z <= 1e-7f) return op_sample_none();
        event = 2;
    } else {
        const float eta = op_modulated_ior(m);
        const float mep = entering ? eta : 1.0f / eta;
        float ep = 1.0f;
        v3 wm = mk3(0.0f, 0.0f, 1.0f);
        if (!smooth) wm = op_sample_vndf(op_spec_rough(m), wo, u1, u2);
        if (!op_refract(wo, wm, mep, &ep, &wi) || (!smooth && wi.z >= -1e-7f)) return op_sample_none();
        if (smooth) {
            const v3 t = op_clamp_unit(sub3(op_vec(1.0f), op_dielectric_reflectance(m, wo.z, entering)));
            return op_delta(&f, wi, scl3(t, 1.0f / (ep * ep)), p.transmission, 3, 1);
        }
        event = 3;
    }
    return op_finish(m, n, &f, won, wi, event, thr);
}

static op_sample op_sample_full(const rto_openpbr* m, v3 n, v3 wo_world, float u_lobe, float u1, float u2, v3 thr) {
    const v3 won = op_normalize(wo_world);
    if (lensq3(won) <= 0.0f) return op_sample_none();
    op_frame f = op_make_frame(n);
    const int entering = m->geometry_thin_walled != 0 || dot3(f.n, won) >= 0.0f;
    f = op_face(f, won);
    const v3 wo = op_to_local(&f, won);
    if (wo.z <= 1e-7f) return op_sample_none();
    const float opacity = op_sat(m->geometry_opacity);
    const float pass = 1.0f - opacity;
    u_lobe = op_sat(u_lobe);
    if (u_lobe < pass) {
        op_sample s = op_sample_none();
        s.wi = neg3(won);
        s.weight = op_vec(1.0f);
        s.discrete_pdf = pass;
        s.event = 5;
        s.delta = 1;
        s.valid = 1;
        return s;
    }
    if (opacity <= 0.0f) return op_sample_none();
    const float choice = op_clamp((u_lobe - pass) / opacity, 0.0f, 0.99999994f);
    if (!entering && op_ss_random_walk(m)) {
        /* leaving the subsurface medium through its boundary */
        const rto_openpbr b = op_boundary_material(m);
        op_sample s = op_sample_boundary(&b, n, won, choice, u1, u2, thr);
        if (!s.valid) return s;
        const v3 sc = op_substrate_scale(m, wo.z, dot3(f.n, s.wi), 0);
        s.value = mul3(s.value, scl3(sc, opacity));
        s.weight = mul3(s.weight, sc);
        s.pdf *= opacity;
        s.discrete_pdf *= opacity;
        if (s.event == 3) s.event = 9;
        return s;
    }
    const int exterior = entering;
    const op_probs p = op_lobe_probs(m, wo.z, exterior);
    if (p.diffuse + p.subsurface + p.reflection + p.transmission + p.coat + p.fuzz <= 0.0f)
        return op_sample_none();
    const float fuzz_end = p.fuzz;
    const float coat_end = fuzz_end + p.coat;
    const float diffuse_end = coat_end + p.diffuse;
    const float ss_end = diffuse_end + p.subsurface;
    const float refl_end = ss_end + p.reflection;
    v3 wi;
    int event;
    if (choice < fuzz_end) {
        wi = op_sample_fuzz(m, wo, u1, u2);
        if (wi.z <= 1e-7f) return op_sample_none();
        event = 7;
    } else if (choice < coat_end) {
        const op_rough cr = op_coat_rough(m);
        if (op_smooth(cr)) {
            wi = mk3(-wo.x, -wo.y, wo.z);
            v3 coeff = op_vec(op_sat(m->coat_weight) * op_coat_fresnel(m, wo.z) * opacity);
            coeff = scl3(coeff, op_fuzz_base_scale(m, wo.z, wi.z, exterior));
            return op_delta(&f, wi, coeff, opacity * p.coat, 6, 1);
        }
        const v3 wm = op_sample_vndf(cr, wo, u1, u2);
        wi = op_reflect(wo, wm);
        if (wi.z <= 1e-7f) return op_sample_none();
        event = 6;
    } else if (choice < diffuse_end) {
        wi = op_cos_hemisphere(u1, u2);
        event = 1;
    } else if (choice < ss_end) {
        if (op_ss_thin_walled(m)) {
            const float cc = op_clamp((choice - diffuse_end) / op_max(p.subsurface, 1e-8f), 0.0f, 0.99999994f);
            const float rp = op_ss_thin_prob(m, 0);
            wi = op_cos_hemisphere(u1, u2);
            if (cc < rp) {
                event = 1;
            } else {
                wi.z = -wi.z;
                event = 4;
            }
        } else {
            /* entering the subsurface medium through its boundary */
            const rto_openpbr b = op_boundary_material(m);
            const op_probs bp = op_lobe_probs(&b, wo.z, 0);
            if (bp.transmission <= 0.0f || p.subsurface <= 0.0f) return op_sample_none();
            const float cc = op_clamp((choice - diffuse_end) / p.subsurface, 0.0f, 0.99999994f);
            const float bc = bp.reflection + cc * bp.transmission;
            op_sample s = op_sample_boundary(&b, n, won, bc, u1, u2, thr);
            if (!s.valid || s.event != 3) return op_sample_none();
            const float ms = (1.0f - op_sat(m->base_metalness)) * (1.0f - op_sat(m->transmission_weight))
                             * op_sat(m->subsurface_weight);
            const v3 ls = op_substrate_scale(m, wo.z, dot3(f.n, s.wi), 1);
            const v3 ws = scl3(ls, ms * bp.transmission / op_max(p.subsurface, 1e-8f));
            s.value = mul3(s.value, scl3(ls, ms * opacity));
            s.weight = mul3(s.weight, ws);
            s.pdf *= opacity * p.subsurface / bp.transmission;
            s.discrete_pdf *= opacity * p.subsurface / bp.transmission;
            s.event = 8;
            return s;
        }
    } else if (choice < refl_end) {
        if (op_smooth_m(m)) {
            wi = mk3(-wo.x, -wo.y, wo.z);
            v3 coeff = scl3(op_reflection_fresnel(m, wo.z, wo.z, entering), opacity);
            coeff = mul3(coeff, op_substrate_scale(m, wo.z, wi.z, exterior));
            return op_delta(&f, wi, coeff, opacity * p.reflection, 2, 1);
        }
        const v3 wm = op_sample_vndf(op_spec_rough(m), wo, u1, u2);
        wi = op_reflect(wo, wm);
        if (wi.z <= 1e-7f) return op_sample_none();
        event = 2;
    } else if (m->geometry_thin_walled != 0) {
        wi = neg3(wo);
        const v3 t = op_clamp_unit(sub3(op_vec(1.0f), op_dielectric_reflectance(m, wo.z, 1)));
        v3 coeff = scl3(mul3(op_surface_transmission_color(m), t),
                        opacity * (1.0f - op_sat(m->base_metalness)) * op_sat(m->transmission_weight));
        coeff = mul3(coeff, op_substrate_scale(m, wo.z, wi.z, exterior));
        return op_delta(&f, wi, coeff, opacity * p.transmission, 4, 0);
    } else if (op_smooth_m(m)) {
        if (op_dispersion(m)) {
            const float cu = op_clamp((choice - refl_end) / op_max(p.transmission, 1e-8f), 0.0f, 0.99999994f);
            const int ch = op_sample_channel(thr, cu);
            const v3 iors = op_dispersion_ior(m);
            const float mep = entering ? op_comp(iors, ch) : 1.0f / op_comp(iors, ch);
            float rep = 1.0f;
            if (!op_refract(wo, mk3(0.0f, 0.0f, 1.0f), mep, &rep, &wi)) return op_sample_none();
            const v3 t = op_clamp_unit(sub3(op_vec(1.0f), op_dielectric_reflectance(m, wo.z, entering)));
            const v3 tc = op_surface_transmission_color(m);
            v3 coeff = op_vec(0.0f);
            op_set_comp(&coeff, ch, op_comp(tc, ch) * op_comp(t, ch) * opacity
                                        * (1.0f - op_sat(m->base_metalness)) * op_sat(m->transmission_weight)
                                        / (rep * rep));
            coeff = mul3(coeff, op_substrate_scale(m, wo.z, wi.z, exterior));
            const float cp = op_comp(op_channel_probs(thr), ch);
            return op_delta(&f, wi, coeff, opacity * p.transmission * cp, 3, 1);
        }
        const float mep = entering ? op_modulated_ior(m) : 1.0f / op_modulated_ior(m);
        float rep = 1.0f;
        if (!op_refract(wo, mk3(0.0f, 0.0f, 1.0f), mep, &rep, &wi)) return op_sample_none();
        const v3 t = op_clamp_unit(sub3(op_vec(1.0f), op_dielectric_reflectance(m, wo.z, entering)));
        v3 coeff = scl3(mul3(op_surface_transmission_color(m), t),
                        opacity * (1.0f - op_sat(m->base_metalness)) * op_sat(m->transmission_weight)
                            / (rep * rep));
        coeff = mul3(coeff, op_substrate_scale(m, wo.z, wi.z, exterior));
        return op_delta(&f, wi, coeff, opacity * p.transmission, 3, 1);
    } else {
        const v3 wm = op_sample_vndf(op_spec_rough(m), wo, u1, u2);
        float mep = entering ? op_modulated_ior(m) : 1.0f / op_modulated_ior(m);
        if (op_dispersion(m)) {
            const float cu = op_clamp((choice - refl_end) / op_max(p.transmission, 1e-8f), 0.0f, 0.99999994f);
            const int ch = op_sample_channel(thr, cu);
            const v3 iors = op_dispersion_ior(m);
            mep = entering ? op_comp(iors, ch) : 1.0f / op_comp(iors, ch);
        }
        float ep = 1.0f;
        if (!op_refract(wo, wm, mep, &ep, &wi) || wi.z >= -1e-7f) return op_sample_none();
        event = 3;
    }
    return op_finish(m, n, &f, won, wi, event, thr);
}

void om_openpbr_eval(const rto_openpbr* m, v3 n, v3 wo, v3 wi, v3 thr, v3* value, float* pdf) {
    op_eval(m, n, wo, wi, thr, value, pdf);
}

void om_openpbr_sample(const rto_openpbr* m, v3 n, v3 wo, float u_lobe, float u1, float u2, v3 thr,
    rto_openpbr_sample* out) {
    const op_sample s = op_sample_full(m, n, wo, u_lobe, u1, u2, thr);
    out->wi[0] = s.wi.x; out->wi[1] = s.wi.y; out->wi[2] = s.wi.z;
    out->weight[0] = s.weight.x; out->weight[1] = s.weight.y; out->weight[2] = s.weight.z;
    out->pdf = s.pdf;
    out->delta = s.delta;
    out->valid = s.valid;
    out->event = s.event;
}

int om_openpbr_random_walk(const rto_openpbr* m) { return op_ss_random_walk(m); }

void om_openpbr_segment(const rto_openpbr* m, float boundary, float u, v3 thr, float out[5]) {
    const op_medium md = op_ss_medium(m);
    op_ss_segment(&md, boundary, u, thr, out);
}

v3 om_openpbr_hg(const rto_openpbr* m, v3 incident, float u1, float u2) {
    const op_medium md = op_ss_medium(m);
    return op_sample_hg(incident, md.g, u1, u2);
}

/* ---- exports */
void rto_openpbr_evaluate(const rto_openpbr* m, const float n[3], const float wo[3],
    const float wi[3], rto_openpbr_eval* out) {
    v3 val;
    op_eval(m, ld3(n), ld3(wo), ld3(wi), op_vec(1.0f), &val, &out->pdf);
    out->value[0] = val.x; out->value[1] = val.y; out->value[2] = val.z;
}

void rto_openpbr_evaluate_ctx(const rto_openpbr* m, const float n[3], const float wo[3],
    const float wi[3], const float thr[3], rto_openpbr_eval* out) {
    v3 val;
    op_eval(m, ld3(n), ld3(wo), ld3(wi), ld3(thr), &val, &out->pdf);
    out->value[0] = val.x; out->value[1] = val.y; out->value[2] = val.z;
}

void rto_openpbr_sample_fn(const rto_openpbr* m, const float n[3], const float wo[3], float u_lobe,
    float u1, float u2, rto_openpbr_sample* out) {
    om_openpbr_sample(m, ld3(n), ld3(wo), u_lobe, u1, u2, op_vec(1.0f), out);
}

void rto_openpbr_sample_ctx(const rto_openpbr* m, const float n[3], const float wo[3], float u_lobe,
    float u1, float u2, const float thr[3], rto_openpbr_sample* out) {
    om_openpbr_sample(m, ld3(n), ld3(wo), u_lobe, u1, u2, ld3(thr), out);
}

void rto_openpbr_emission(const rto_openpbr* m, float out[3]) {
    const v3 e = om_openpbr_emission(m);
    out[0] = e.x; out[1] = e.y; out[2] = e.z;
}

void rto_openpbr_subsurface_medium(const rto_openpbr* m, float out[8]) {
    const op_medium md = op_ss_medium(m);
    out[0] = md.extinction.x; out[1] = md.extinction.y; out[2] = md.extinction.z;
    out[3] = md.albedo.x; out[4] = md.albedo.y; out[5] = md.albedo.z;
    out[6] = md.g;
    out[7] = (float)md.active;
}

void rto_openpbr_subsurface_segment(const rto_openpbr* m, float boundary, float u,
    const float thr[3], float out[5]) {
    om_openpbr_segment(m, boundary, u, ld3(thr), out);
}

void rto_openpbr_hg(const float incident[3], float g, float u1, float u2, float out[3]) {
    const v3 d = op_sample_hg(ld3(incident), g, u1, u2);
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
}
