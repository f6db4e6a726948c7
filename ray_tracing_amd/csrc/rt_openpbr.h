// rt_openpbr.h -- the OpenPBR surface for the shade stage (and for the host
// light-CDF builder, hence __host__ __device__).
//
// Follows src/common/openpbr_core.h lobe for lobe:
//   transport context (RGB wavelengths, path throughput)        :27-31,255-270
//   transmission colour :360-366; subsurface medium, free-flight segments,
//     Henyey-Greenstein phase                                     :387-519,576-590
//   Fresnel (scalar, polarized, real-IOR phase)                   :592-660
//   thin-film Airy reflectance (Belcour-Barla sensitivity)        :662-813
//   modulated IOR, dispersion                                     :815-861
//   anisotropic GGX with energy compensation, VNDF, refraction    :863-1042
//   coat :1044-1107, fuzz (LTC sheen) :1109-1209, layering :1211-1298
//   lobe probabilities :1306-1387; Fujii Oren-Nayar diffuse :1389-1473
//   dielectric / metal reflectance with thin film :1475-1677
//   evaluate :1679-1860, emission :1868-1883, subsurface boundary
//   :1891-1993, sample :1995-2297
// A material is read through a view type (accessor per field): the plain
// view reads the packed record in memory, and wrapper views replace fields
// (the subsurface boundary's clear dielectric, the thin-walled subsurface's
// diffuse layer) with compile-time constants, so the omitted terms fold away
// instead of copying the 204-byte record into registers. Sampled azimuths
// use sincos_turn instead of cosf/sinf(2*pi*u).
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

// openpbr_sample_subsurface_segment
struct OpenPbrSegment {
    f3 weight;
    float distance;
    bool scattered;
};

// scatter events (OpenPbrScatterEvent, openpbr_core.h:131-142)
enum : int { kPbrNone = 0, kPbrDiffuse = 1, kPbrGlossyReflection = 2, kPbrGlossyTransmission = 3,
    kPbrThinTransmission = 4, kPbrPassthrough = 5, kPbrCoat = 6, kPbrFuzz = 7,
    kPbrSubsurfaceEntry = 8, kPbrSubsurfaceExit = 9 };

// openpbr_srgb_channel_to_linear / openpbr_source_to_linear
// (openpbr_core.h:322-333): a texture value in the sRGB-encoded colour space
// is decoded, the other spaces are taken as linear
RTK_HD float srgb_channel_to_linear(float v) {
    return v <= 0.04045f ? v / 12.92f : powf((v + 0.055f) / 1.055f, 2.4f);
}
RTK_HD f3 source_to_linear(f3 c, int space) {
    if (space != RT_COLOR_SRGB_TEXTURE) return c;
    return f3{srgb_channel_to_linear(c.x), srgb_channel_to_linear(c.y), srgb_channel_to_linear(c.z)};
}

namespace opbr {

constexpr float kPiF = 3.14159265358979323846f;
constexpr float kInvPi = 1.0f / kPiF;
constexpr float kFraunhoferC = 656.3f;
constexpr float kFraunhoferD = 587.6f;
constexpr float kFraunhoferF = 486.1f;

struct Rough {
    float ax, ay;
};
struct Frame {
    f3 t, b, n;
};
struct Probs {
    float diffuse, subsurface, reflection, transmission, coat, fuzz;
};
struct Pol {
    float p, s;
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
RTK_HD float sq(float x) { return x * x; }
RTK_HD float safe_sqrt(float x) { return sqrtf(omax(0.0f, x)); }
RTK_HD f3 clamp_unit(f3 a) { return v(sat(a.x), sat(a.y), sat(a.z)); }
RTK_HD f3 clamp_nonneg(f3 a) { return v(omax(0.0f, a.x), omax(0.0f, a.y), omax(0.0f, a.z)); }
RTK_HD f3 odiv(f3 a, float s) {
    const float inv = 1.0f / s;
    return scl(a, inv);
}
RTK_HD f3 lerp(f3 a, f3 b, float t) { return add(scl(a, 1.0f - t), scl(b, t)); }
RTK_HD f3 recip(f3 a) {
    return v(1.0f / omax(a.x, 1e-6f), 1.0f / omax(a.y, 1e-6f), 1.0f / omax(a.z, 1e-6f));
}
RTK_HD float comp(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
RTK_HD void set_comp(f3& a, int i, float x) {
    if (i == 0) a.x = x;
    else if (i == 1) a.y = x;
    else a.z = x;
}
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
RTK_HD f3 sqrt3(f3 a) { return v(safe_sqrt(a.x), safe_sqrt(a.y), safe_sqrt(a.z)); }
RTK_HD f3 pow3(f3 a, float e) {
    return v(powf(omax(a.x, 0.0f), e), powf(omax(a.y, 0.0f), e), powf(omax(a.z, 0.0f), e));
}

// ---- material views
#define RTK_PBR_SCALARS(X)                                                                         \
    X(base_weight) X(base_diffuse_roughness) X(base_metalness) X(specular_weight)                  \
    X(specular_roughness) X(specular_ior) X(specular_roughness_anisotropy) X(transmission_weight)  \
    X(transmission_depth) X(transmission_dispersion_scale) X(transmission_dispersion_abbe_number)  \
    X(subsurface_weight) X(subsurface_radius) X(subsurface_scatter_anisotropy) X(fuzz_weight)      \
    X(fuzz_roughness) X(coat_weight) X(coat_roughness) X(coat_roughness_anisotropy) X(coat_ior)    \
    X(coat_darkening) X(thin_film_weight) X(thin_film_thickness) X(thin_film_ior)                  \
    X(emission_luminance) X(geometry_opacity)
#define RTK_PBR_VECTORS(X)                                                                         \
    X(base_color) X(specular_color) X(transmission_color) X(subsurface_color)                      \
    X(subsurface_radius_scale) X(fuzz_color) X(coat_color) X(emission_color)

// the packed record in memory
struct MatMem {
    static constexpr bool kFull = true;   // the lobes beyond the opaque core may be present
    const rt_openpbr_material* p;
#define RTK_S(f) RTK_HD float f() const { return p->f; }
#define RTK_V(f) RTK_HD f3 f() const { return ld(p->f); }
    RTK_PBR_SCALARS(RTK_S)
    RTK_PBR_VECTORS(RTK_V)
#undef RTK_S
#undef RTK_V
    RTK_HD int geometry_thin_walled() const { return p->geometry_thin_walled; }
};

// a material of the opaque core: base, specular, metal, coat and emission.
// The other lobes' inputs read as zero at compile time, so their code folds
// away (a register budget the shade and ReSTIR kernels can keep), and for a
// record whose other inputs are zero every value is the same as MatMem's.
struct MatCore : MatMem {
    static constexpr bool kFull = false;
    RTK_HD explicit MatCore(const rt_openpbr_material* m) : MatMem{m} {}
    RTK_HD float specular_roughness_anisotropy() const { return 0.0f; }
    RTK_HD float transmission_weight() const { return 0.0f; }
    RTK_HD float transmission_dispersion_scale() const { return 0.0f; }
    RTK_HD float subsurface_weight() const { return 0.0f; }
    RTK_HD float fuzz_weight() const { return 0.0f; }
    RTK_HD float coat_roughness_anisotropy() const { return 0.0f; }
    RTK_HD float thin_film_weight() const { return 0.0f; }
    RTK_HD int geometry_thin_walled() const { return 0; }
};
// whether a record needs more than the core
RTK_HD bool needs_full(const rt_openpbr_material& m) {
    return m.transmission_weight > 0.0f || m.subsurface_weight > 0.0f || m.fuzz_weight > 0.0f
           || m.thin_film_weight > 0.0f || m.transmission_dispersion_scale > 0.0f
           || m.specular_roughness_anisotropy != 0.0f || m.coat_roughness_anisotropy != 0.0f
           || m.geometry_thin_walled != 0;
}

// a material with the inputs that textures may drive (base, specular and
// transmission colour, metalness, roughness, emission colour) held as values:
// the record's own, or a texture's at the hit (the per-hit parameter copy of
// evaluate_openpbr_scattering_material / populate_material,
// programs.cu:944-959,995-1008)
template <class M>
struct MatTex : M {
    f3 bc, sc, tc, ec;
    float metal, rough;
    RTK_HD explicit MatTex(const M& m)
        : M(m), bc(m.base_color()), sc(m.specular_color()), tc(m.transmission_color()),
          ec(m.emission_color()), metal(m.base_metalness()), rough(m.specular_roughness()) {}
    RTK_HD f3 base_color() const { return bc; }
    RTK_HD f3 specular_color() const { return sc; }
    RTK_HD f3 transmission_color() const { return tc; }
    RTK_HD f3 emission_color() const { return ec; }
    RTK_HD float base_metalness() const { return metal; }
    RTK_HD float specular_roughness() const { return rough; }
};

// the medium of a path inside a subsurface random walk: the walk started
// only where the metalness at the entry point was below one, so the
// metalness gate of ss_random_walk holds inside whatever a texture gave there
// (the medium's other inputs are never texture-driven)
template <class M>
struct MatInside : M {
    RTK_HD explicit MatInside(const M& m) : M(m) {}
    RTK_HD float base_metalness() const { return 0.0f; }
};

// openpbr_subsurface_boundary_material (openpbr_core.h:1891-1905): a clear
// dielectric with the material's IOR, roughness and thin film
template <class M>
struct MatBoundary : M {
    RTK_HD explicit MatBoundary(const M& m) : M(m) {}
    RTK_HD float base_weight() const { return 0.0f; }
    RTK_HD float base_metalness() const { return 0.0f; }
    RTK_HD float transmission_weight() const { return 1.0f; }
    RTK_HD f3 transmission_color() const { return vs(1.0f); }
    RTK_HD float transmission_depth() const { return 0.0f; }
    RTK_HD float transmission_dispersion_scale() const { return 0.0f; }
    RTK_HD float subsurface_weight() const { return 0.0f; }
    RTK_HD float fuzz_weight() const { return 0.0f; }
    RTK_HD float coat_weight() const { return 0.0f; }
    RTK_HD float geometry_opacity() const { return 1.0f; }
};

// the diffuse layer of thin-walled subsurface (openpbr_core.h:1461-1473)
template <class M>
struct MatThinDiffuse : M {
    float weight;
    RTK_HD MatThinDiffuse(const M& m, float w) : M(m), weight(w) {}
    RTK_HD float base_weight() const { return weight; }
    RTK_HD f3 base_color() const { return clamp_unit(M::subsurface_color()); }
    RTK_HD float subsurface_weight() const { return 0.0f; }
};

// ---- transport context
RTK_HD f3 channel_probs(f3 thr) {
    const f3 t = clamp_nonneg(thr);
    const float total = t.x + t.y + t.z;
    return total > 1e-8f ? odiv(t, total) : vs(1.0f / 3.0f);
}
RTK_HD int sample_channel(f3 thr, float u) {
    const f3 p = channel_probs(thr);
    const float c = oclamp(u, 0.0f, 0.99999994f);
    return c < p.x ? 0 : (c < p.x + p.y ? 1 : 2);
}

template <class M>
RTK_HD f3 surface_transmission_color(const M& m) {
    if (m.geometry_thin_walled() != 0 || m.transmission_depth() <= 0.0f)
        return clamp_unit(m.transmission_color());
    return vs(1.0f);
}

// ---- subsurface random walk
RTK_HD float ss_albedo(float observed, float anisotropy) {
    const float color = sat(observed);
    const float g = oclamp(anisotropy, -0.999f, 0.999f);
    const float s = omax(0.0f, 4.09712f + 4.20863f * color
                                   - sqrtf(omax(0.0f, 9.59217f + 41.6808f * color
                                                          + 17.7126f * color * color)));
    const float s2 = s * s;
    return sat((1.0f - s2) / omax(1.0f - g * s2, 1e-6f));
}
template <class M>
RTK_HD bool ss_finite_mfp(const M& m) {
    const float radius = omax(0.0f, m.subsurface_radius());
    const f3 sc = clamp_nonneg(m.subsurface_radius_scale());
    return radius * omax(sc.x, omax(sc.y, sc.z)) > 1e-6f;
}
template <class M>
RTK_HD bool ss_random_walk(const M& m) {
    return m.geometry_thin_walled() == 0 && sat(m.subsurface_weight()) > 0.0f
           && sat(m.base_metalness()) < 1.0f && sat(m.transmission_weight()) < 1.0f
           && ss_finite_mfp(m);
}
template <class M>
RTK_HD bool ss_thin_walled(const M& m) {
    return m.geometry_thin_walled() != 0 && sat(m.subsurface_weight()) > 0.0f;
}
template <class M>
RTK_HD float ss_thin_prob(const M& m, bool transmission) {
    const float g = oclamp(m.subsurface_scatter_anisotropy(), -1.0f, 1.0f);
    return 0.5f * (1.0f + (transmission ? g : -g));
}

struct Medium {
    f3 extinction, albedo;
    float g;
    bool active;
};
template <class M>
RTK_HD Medium ss_medium(const M& m) {
    Medium md{vs(0.0f), vs(0.0f), 0.0f, false};
    if (!ss_random_walk(m)) return md;
    const float radius = omax(0.0f, m.subsurface_radius());
    const f3 sc = clamp_nonneg(m.subsurface_radius_scale());
    const f3 mfp = v(omax(radius * sc.x, 1e-6f), omax(radius * sc.y, 1e-6f), omax(radius * sc.z, 1e-6f));
    md.extinction = recip(mfp);
    const f3 color = clamp_unit(m.subsurface_color());
    const float an = m.subsurface_scatter_anisotropy();
    md.albedo = v(ss_albedo(color.x, an), ss_albedo(color.y, an), ss_albedo(color.z, an));
    md.g = oclamp(an, -0.999f, 0.999f);
    md.active = true;
    return md;
}
RTK_HD f3 ss_channel_probs(const Medium& md, f3 thr) {
    const f3 c = mul(clamp_nonneg(thr), md.albedo);
    const float total = c.x + c.y + c.z;
    return total > 1e-12f ? odiv(c, total) : vs(1.0f / 3.0f);
}
RTK_HD f3 ss_transmittance(const Medium& md, float d) {
    const float sd = omax(0.0f, d);
    return v(expf(-md.extinction.x * sd), expf(-md.extinction.y * sd), expf(-md.extinction.z * sd));
}
RTK_HD OpenPbrSegment ss_segment(const Medium& md, float boundary, float u, f3 thr) {
    OpenPbrSegment r{vs(1.0f), omax(0.0f, boundary), false};
    if (!md.active || r.distance <= 0.0f) return r;
    const f3 p = ss_channel_probs(md, thr);
    const float choice = oclamp(u, 0.0f, 0.99999994f);
    int channel = 2;
    float cu = 0.0f, cdf = 0.0f;
    for (int i = 0; i < 3; ++i) {
        const float pi = comp(p, i);
        if (choice < cdf + pi || i == 2) {
            channel = i;
            cu = (choice - cdf) / omax(pi, 1e-8f);
            break;
        }
        cdf += pi;
    }
    const float ext = comp(md.extinction, channel);
    const float sampled = -logf(omax(1.0f - sat(cu), 1e-7f)) / omax(ext, 1e-6f);
    r.scattered = sampled < r.distance;
    r.distance = r.scattered ? sampled : r.distance;
    const f3 tr = ss_transmittance(md, r.distance);
    if (r.scattered) {
        const f3 density = mul(md.extinction, tr);
        const float mix = dot(p, density);
        r.weight = odiv(mul(density, md.albedo), omax(mix, 1e-12f));
    } else {
        const float surv = dot(p, tr);
        r.weight = odiv(tr, omax(surv, 1e-12f));
    }
    return r;
}

// ---- frames
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

RTK_HD f3 sample_hg(f3 incident, float anisotropy, float u1, float u2) {
    const float g = oclamp(anisotropy, -0.999f, 0.999f);
    float c = 1.0f - 2.0f * sat(u1);
    if (fabsf(g) > 1e-3f) {
        const float ratio = (1.0f - g * g) / omax(1.0f - g + 2.0f * g * sat(u1), 1e-6f);
        c = oclamp((1.0f + g * g - ratio * ratio) / (2.0f * g), -1.0f, 1.0f);
    }
    const float s = safe_sqrt(1.0f - c * c);
    float sp, cp;
    sincos_turn(sat(u2), &sp, &cp);
    const f3 local = v(s * cp, s * sp, c);
    const Frame f = make_frame(onormalize(incident));
    return onormalize(to_world(f, local));
}

// ---- Fresnel
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
RTK_HD Pol fresnel_polarized(float cosine, float eta) {
    const float c2 = sq(sat(cosine));
    const float s2 = 1.0f - c2;
    const float t0 = omax(eta * eta - s2, 0.0f);
    const float t1 = t0 + c2;
    const float t2 = 2.0f * sqrtf(t0) * sat(cosine);
    const float rs = (t1 - t2) / omax(t1 + t2, 1e-8f);
    const float t3 = c2 * t0 + s2 * s2;
    const float t4 = t2 * s2;
    const float rp = rs * (t3 - t4) / omax(t3 + t4, 1e-8f);
    return Pol{rp, rs};
}
RTK_HD Pol fresnel_real_polarized(float cosine, float rel) {
    const float c2 = sq(sat(cosine));
    const float s2 = 1.0f - c2;
    const float t0 = rel * rel - s2;
    const float mag = fabsf(t0);
    const float t1 = mag + c2;
    const float a = safe_sqrt(0.5f * (mag + t0));
    const float t2 = 2.0f * a * sat(cosine);
    const float rs = (t1 - t2) / omax(t1 + t2, 1e-8f);
    const float t3 = c2 * mag + s2 * s2;
    const float t4 = t2 * s2;
    const float rp = rs * (t3 - t4) / omax(t3 + t4, 1e-8f);
    return Pol{rp, rs};
}
RTK_HD Pol real_phase(float cosine, float inc, float trans) {
    const float s2 = 1.0f - sq(sat(cosine));
    const float t2 = trans * trans, i2 = inc * inc;
    const float a = t2 - i2 * s2;
    const float u = safe_sqrt(0.5f * (a + fabsf(a)));
    const float sd = u * u - sq(inc * cosine);
    const float pd = sq(t2 * cosine) - i2 * u * u;
    return Pol{pd < 0.0f ? kPiF : 0.0f, sd < 0.0f ? kPiF : 0.0f};
}

// ---- thin film
RTK_HD f3 film_sensitivity(float opd, f3 shift) {
    const float phase = 2.0f * kPiF * opd;
    const f3 val = v(5.4856e-13f, 4.4201e-13f, 5.2481e-13f);
    const f3 pos = v(1.6810e6f, 1.7953e6f, 2.2084e6f);
    const f3 var = v(4.3278e9f, 9.3046e9f, 6.6121e9f);
    f3 xyz = v(val.x * sqrtf(2.0f * kPiF * var.x) * cosf(pos.x * phase + shift.x) * expf(-var.x * phase * phase),
               val.y * sqrtf(2.0f * kPiF * var.y) * cosf(pos.y * phase + shift.y) * expf(-var.y * phase * phase),
               val.z * sqrtf(2.0f * kPiF * var.z) * cosf(pos.z * phase + shift.z) * expf(-var.z * phase * phase));
    xyz.x += 9.7470e-14f * sqrtf(2.0f * kPiF * 4.5282e9f) * cosf(2.2399e6f * phase + shift.x)
             * expf(-4.5282e9f * phase * phase);
    return odiv(xyz, 1.0685e-7f);
}
RTK_HD f3 xyz_to_rgb(f3 c) {
    return v(2.3706743f * c.x - 0.9000405f * c.y - 0.4706338f * c.z,
             -0.5138850f * c.x + 1.4253036f * c.y + 0.0885814f * c.z,
             0.0052982f * c.x - 0.0146949f * c.y + 1.0093968f * c.z);
}
RTK_HD float f0_to_ior(float f0) {
    const float r = sqrtf(oclamp(f0, 0.01f, 0.99f));
    return (1.0f + r) / omax(1.0f - r, 1e-5f);
}
RTK_HD float film_cos_t(float cosine, float inc, float film) {
    const float ratio = inc / omax(film, 1e-4f);
    const float s2 = (1.0f - sq(sat(cosine))) * ratio * ratio;
    return s2 < 1.0f ? safe_sqrt(1.0f - s2) : 0.0f;
}
// one polarization's Airy sum with its first two interference orders
RTK_HD f3 airy_orders(float r12, float t121, f3 r23, f3 ph23, float ph21, f3 r123, float opd) {
    const f3 tr = v(t121 * t121 * r23.x / omax(1.0f - r12 * r23.x, 1e-6f),
                    t121 * t121 * r23.y / omax(1.0f - r12 * r23.y, 1e-6f),
                    t121 * t121 * r23.z / omax(1.0f - r12 * r23.z, 1e-6f));
    f3 result = add(vs(r12), tr);
    f3 coeff = sub(tr, vs(t121));
    for (int order = 1; order <= 2; ++order) {
        coeff = mul(coeff, r123);
        const f3 shift = scl(add(ph23, vs(ph21)), (float)order);
        result = add(result, mul(coeff, scl(film_sensitivity((float)order * opd, shift), 2.0f)));
    }
    return result;
}
RTK_HD f3 film_airy(float cosine, float thick_nm, float inc_ior, float film_ior, f3 sub_ior,
    f3 gen_refl, bool gen_schlick) {
    const float eta1 = omax(inc_ior, 1e-4f);
    const float eta2 = omax(film_ior, 1e-4f);
    const float ct = sat(fabsf(cosine));
    const float ctt = film_cos_t(ct, eta1, eta2);
    if (ctt <= 0.0f) return vs(1.0f);
    const Pol r12 = fresnel_polarized(ct, eta2 / eta1);
    const Pol t121{1.0f - r12.p, 1.0f - r12.s};
    f3 r23p = vs(0.0f), r23s = vs(0.0f), ph23p = vs(0.0f), ph23s = vs(0.0f);
    for (int c = 0; c < 3; ++c) {
        const float eta3 = comp(sub_ior, c);
        if (gen_schlick) {
            const float refl = 0.5f * sat(comp(gen_refl, c));
            const float ph = eta3 < eta2 ? kPiF : 0.0f;
            set_comp(r23p, c, refl);
            set_comp(r23s, c, refl);
            set_comp(ph23p, c, ph);
            set_comp(ph23s, c, ph);
        } else {
            const Pol refl = fresnel_real_polarized(ctt, eta3 / eta2);
            const Pol ph = real_phase(ctt, eta2, eta3);
            set_comp(r23p, c, refl.p);
            set_comp(r23s, c, refl.s);
            set_comp(ph23p, c, ph.p);
            set_comp(ph23s, c, ph.s);
        }
    }
    const float brewster = eta1 / sqrtf(eta1 * eta1 + eta2 * eta2);
    const Pol ph21{ct < brewster ? 0.0f : kPiF, kPiF};
    const f3 r123p = v(sqrtf(omax(r12.p * r23p.x, 0.0f)), sqrtf(omax(r12.p * r23p.y, 0.0f)),
                       sqrtf(omax(r12.p * r23p.z, 0.0f)));
    const f3 r123s = v(sqrtf(omax(r12.s * r23s.x, 0.0f)), sqrtf(omax(r12.s * r23s.y, 0.0f)),
                       sqrtf(omax(r12.s * r23s.z, 0.0f)));
    const float opd = 2.0f * eta2 * ctt * thick_nm * 1e-9f;
    f3 result = airy_orders(r12.p, t121.p, r23p, ph23p, ph21.p, r123p, opd);
    result = add(result, airy_orders(r12.s, t121.s, r23s, ph23s, ph21.s, r123s, opd));
    return clamp_unit(xyz_to_rgb(scl(result, 0.5f)));
}

// ---- IOR and dispersion
template <class M>
RTK_HD float modulated_ior(const M& m) {
    const float eta = omax(m.specular_ior(), 1e-4f);
    const float ratio = (eta - 1.0f) / (eta + 1.0f);
    const float f0 = oclamp(omax(0.0f, m.specular_weight()) * ratio * ratio, 0.0f, 0.9999f);
    const float sign = eta < 1.0f ? -1.0f : (eta > 1.0f ? 1.0f : 0.0f);
    const float eps = sign * sqrtf(f0);
    return omax((1.0f + eps) / omax(1.0f - eps, 1e-4f), 1e-4f);
}
template <class M>
RTK_HD bool dispersion(const M& m) {
    return m.transmission_dispersion_scale() > 0.0f && sat(m.base_metalness()) < 1.0f;
}
RTK_HD float dispersion_ior1(float ior0, float scale_v, float abbe, float wl) {
    const float scale = sat(scale_v);
    if (scale <= 0.0f || fabsf(ior0 - 1.0f) <= 1e-8f) return ior0;
    const bool rcp = ior0 < 1.0f;
    const float ior = rcp ? 1.0f / omax(ior0, 1e-6f) : ior0;
    const float vd = omax(abbe, 1e-4f) / scale;
    const float inv_f2 = 1.0f / sq(kFraunhoferF);
    const float inv_c2 = 1.0f / sq(kFraunhoferC);
    const float b = (ior - 1.0f) / (vd * (inv_f2 - inv_c2));
    const float a = ior - b / sq(kFraunhoferD);
    const float adj = a + b / sq(omax(wl, 1e-4f));
    return rcp ? 1.0f / omax(adj, 1e-6f) : adj;
}
template <class M>
RTK_HD f3 dispersion_ior(const M& m) {
    const float ior = modulated_ior(m);
    if (!dispersion(m)) return vs(ior);
    const float s = m.transmission_dispersion_scale(), ab = m.transmission_dispersion_abbe_number();
    return v(dispersion_ior1(ior, s, ab, kFraunhoferC), dispersion_ior1(ior, s, ab, kFraunhoferD),
             dispersion_ior1(ior, s, ab, kFraunhoferF));
}

// ---- roughness and GGX
RTK_HD Rough roughness(float rv, float av) {
    const float r = sat(rv);
    const float aniso = sat(av);
    const float aspect = 1.0f - aniso;
    const float scale = sqrtf(2.0f / omax(aspect * aspect + 1.0f, 1e-8f));
    const float ax = omax(r * r * scale, 1e-4f);
    return Rough{ax, omax(aspect * ax, 1e-4f)};
}
template <class M>
RTK_HD float eff_spec_roughness(const M& m) {
    const float s = sat(m.specular_roughness());
    const float c = sat(m.coat_roughness());
    const float s2 = s * s, c2 = c * c;
    const float affected = sqrtf(sqrtf(omin(1.0f, s2 * s2 + 2.0f * c2 * c2)));
    const float w = sat(m.coat_weight());
    return s + (affected - s) * w;
}
template <class M>
RTK_HD Rough spec_rough(const M& m) {
    return roughness(eff_spec_roughness(m), m.specular_roughness_anisotropy());
}
template <class M>
RTK_HD Rough coat_rough(const M& m) {
    return roughness(m.coat_roughness(), m.coat_roughness_anisotropy());
}
RTK_HD bool smooth(Rough r) { return omax(r.ax, r.ay) < 1e-3f; }
template <class M>
RTK_HD bool smooth_m(const M& m) { return smooth(spec_rough(m)); }
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
RTK_HD bool refract(f3 wo, f3 wm, float eta, float* eta_path, f3* wi) {
    float ci = dot(wm, wo);
    *eta_path = omax(eta, 1e-4f);
    if (ci < 0.0f) {
        wm = neg(wm);
        ci = -ci;
        *eta_path = 1.0f / *eta_path;
    }
    const float s2i = omax(0.0f, 1.0f - ci * ci);
    const float s2t = s2i / (*eta_path * *eta_path);
    if (s2t >= 1.0f) return false;
    const float ct = safe_sqrt(1.0f - s2t);
    *wi = onormalize(add(scl(wo, -1.0f / *eta_path), scl(wm, ci / *eta_path - ct)));
    return true;
}

// ---- coat
template <class M>
RTK_HD float coat_fresnel(const M& m, float cosine) {
    return fresnel_dielectric(cosine, omax(m.coat_ior(), 1e-4f));
}
template <class M>
RTK_HD float coat_albedo(const M& m, float cosine) {
    const Rough r = coat_rough(m);
    if (smooth(r)) return coat_fresnel(m, cosine);
    const float alpha = avg_alpha(r);
    const float c0 = coat_fresnel(m, 1.0f);
    const float ss = ggx_albedo(cosine, alpha, c0, 1.0f);
    const float white = ggx_albedo(cosine, alpha, 1.0f, 1.0f);
    const float cmp = 1.0f + coat_fresnel(m, cosine) * (1.0f - white) / omax(white, 1e-4f);
    return sat(ss * cmp);
}
template <class M>
RTK_HD f3 coat_passage(const M& m, float cosine) {
    const float presence = sat(m.coat_weight());
    const f3 tint = clamp_unit(m.coat_color());
    if (presence <= 0.0f || cosine <= 0.0f || (tint.x >= 1.0f && tint.y >= 1.0f && tint.z >= 1.0f))
        return vs(1.0f);
    const float eta = omax(m.coat_ior(), 1e-4f);
    const float s2t = omax(0.0f, 1.0f - cosine * cosine) / (eta * eta);
    const float ct = safe_sqrt(1.0f - omin(s2t, 1.0f));
    const float ds = 1.0f / omax(ct, 1e-4f);
    return lerp(vs(1.0f), pow3(sqrt3(tint), ds), presence);
}
template <class M>
RTK_HD f3 coat_darkening(const M& m) {
    const float presence = sat(m.coat_weight()) * sat(m.coat_darkening());
    if (presence <= 0.0f) return vs(1.0f);
    const float eta = omax(m.coat_ior(), 1e-4f);
    const float f0 = coat_fresnel(m, 1.0f);
    const float ir = sat(1.0f - (1.0f - f0) / omax(eta * eta, 1e-6f));
    const f3 da = scl(clamp_unit(m.base_color()), sat(m.base_weight()));
    const f3 ma = scl(da, sat(m.specular_weight()));
    const f3 ba = lerp(da, ma, sat(m.base_metalness()));
    const float num = 1.0f - ir;
    const f3 dk = v(num / omax(1.0f - ba.x * ir, 1e-5f), num / omax(1.0f - ba.y * ir, 1e-5f),
        num / omax(1.0f - ba.z * ir, 1e-5f));
    return lerp(vs(1.0f), clamp_unit(dk), presence);
}

// ---- fuzz
RTK_HD float fuzz_albedo(float cosine, float rv) {
    const float x = sat(cosine);
    const float y = oclamp(rv, 0.01f, 1.0f);
    const float s = y * (0.0206607f + 1.58491f * y) / (0.0379424f + y * (1.32227f + y));
    const float y2 = y * y;
    const float mm = y * (-0.193854f + y * (-1.14885f + y * (1.7932f - 0.95943f * y2))) / (0.046391f + y);
    const float o = y * (0.000654023f + (-0.0207818f + 0.119681f * y) * y) / (1.26264f + y * (-1.92021f + y));
    const float nrm = (x - mm) / omax(s, 1e-5f);
    const float gauss = expf(-0.5f * nrm * nrm) / (omax(s, 1e-5f) * sqrtf(2.0f * kPiF));
    return sat(gauss + o);
}
RTK_HD Frame fuzz_frame(f3 wo) {
    const f3 proj = v(wo.x, wo.y, 0.0f);
    const f3 t = lensq(proj) > 1e-12f ? onormalize(proj) : v(1.0f, 0.0f, 0.0f);
    return Frame{t, v(-t.y, t.x, 0.0f), v(0.0f, 0.0f, 1.0f)};
}
RTK_HD f3 fuzz_ltc(float cosine, float rv) {
    const float x = sat(cosine);
    const float y = oclamp(rv, 0.01f, 1.0f);
    const float ainv = (2.58126f * x + 0.813703f * y) * y / (1.0f + 0.310327f * x * x + 2.60994f * x * y);
    const float binv = safe_sqrt(1.0f - x) * (y - 1.0f) * y * y * y
                       / (0.0000254053f + 1.71228f * x - 1.71506f * x * y + 1.34174f * y * y);
    return v(omax(ainv, 1e-5f), binv, fuzz_albedo(x, y));
}
template <class M>
RTK_HD float fuzz_pdf(const M& m, f3 wo, f3 wi) {
    if (wo.z <= 0.0f || wi.z <= 0.0f) return 0.0f;
    const Frame f = fuzz_frame(wo);
    const f3 al = to_local(f, wi);
    const f3 c = fuzz_ltc(wo.z, m.fuzz_roughness());
    const f3 tr = v(c.x * al.x + c.y * al.z, c.x * al.y, al.z);
    const float l2 = lensq(tr);
    if (l2 <= 1e-16f || tr.z <= 0.0f) return 0.0f;
    const float js = c.x / l2;
    return tr.z * kInvPi * js * js;
}
template <class M>
RTK_HD f3 sample_fuzz(const M& m, f3 wo, float u1, float u2) {
    const float radius = sqrtf(sat(u1));
    float sp, cp;
    sincos_turn(sat(u2), &sp, &cp);
    const f3 cs = v(radius * cp, radius * sp, safe_sqrt(1.0f - radius * radius));
    const f3 c = fuzz_ltc(wo.z, m.fuzz_roughness());
    const float ia = 1.0f / c.x;
    const f3 al = onormalize(v(cs.x * ia - cs.z * c.y * ia, cs.y * ia, cs.z));
    return onormalize(to_world(fuzz_frame(wo), al));
}
template <class M>
RTK_HD f3 fuzz_value(const M& m, f3 wo, f3 wi) {
    const float presence = sat(m.fuzz_weight());
    if (presence <= 0.0f || wo.z <= 0.0f || wi.z <= 0.0f) return vs(0.0f);
    const float refl = fuzz_albedo(wo.z, m.fuzz_roughness());
    const float dens = fuzz_pdf(m, wo, wi);
    return scl(clamp_unit(m.fuzz_color()), presence * refl * dens / omax(wi.z, 1e-6f));
}
template <class M>
RTK_HD float fuzz_base_scale(const M& m, float ocos, float icos, bool exterior) {
    if (!exterior) return 1.0f;
    const float presence = sat(m.fuzz_weight());
    const float out = 1.0f - presence * fuzz_albedo(ocos, m.fuzz_roughness());
    const float in = icos > 0.0f ? 1.0f - presence * fuzz_albedo(icos, m.fuzz_roughness()) : 1.0f;
    return sat(out) * sat(in);
}

// ---- layering
template <class M>
RTK_HD f3 coat_base_scale(const M& m, float ocos, float icos, bool exterior) {
    const float presence = sat(m.coat_weight());
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
template <class M>
RTK_HD f3 substrate_scale(const M& m, float ocos, float icos, bool exterior) {
    return scl(coat_base_scale(m, ocos, icos, exterior), fuzz_base_scale(m, ocos, icos, exterior));
}
template <class M>
RTK_HD f3 coat_value(const M& m, f3 wo, f3 wi) {
    const float presence = sat(m.coat_weight());
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
    const float cmp = energy_comp(wo.z, avg_alpha(r), vs(F)).x;
    return vs(presence * F * cmp * D * G / omax(4.0f * wo.z * wi.z, 1e-8f));
}
template <class M>
RTK_HD float coat_pdf(const M& m, f3 wo, f3 wi) {
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

// ---- base Fresnel terms
template <class M>
RTK_HD float dielectric_fresnel(const M& m, float cosine, bool entering) {
    float eta = modulated_ior(m);
    if (m.geometry_thin_walled() == 0 && !entering) eta = 1.0f / eta;
    float f = fresnel_dielectric(cosine, eta);
    if (m.geometry_thin_walled() != 0 && f < 1.0f) {
        const float t = 1.0f - f;
        f += t * t * f / omax(1.0f - f * f, 1e-6f);
    }
    return sat(f);
}
RTK_HD f3 schlick(f3 c0, f3 c90, float cosine) { return lerp(c0, c90, pow5(1.0f - sat(cosine))); }
template <class M>
RTK_HD f3 metal_fresnel(const M& m, float cosine) {
    const float kEdge = 1.0f / 7.0f;
    const float mu = sat(fabsf(cosine));
    const f3 c0 = scl(clamp_unit(m.base_color()), sat(m.base_weight()));
    const f3 c90 = vs(1.0f);
    const f3 sch = schlick(c0, c90, mu);
    const f3 sch_edge = schlick(c0, c90, kEdge);
    const f3 desired = mul(sch_edge, clamp_unit(m.specular_color()));
    const float omm = 1.0f - mu;
    const float ome = 1.0f - kEdge;
    const float corr = mu * pow5(omm) * omm / (kEdge * pow5(ome) * ome);
    const f3 f82 = sub(sch, scl(sub(sch_edge, desired), corr));
    return scl(clamp_unit(f82), sat(m.specular_weight()));
}
template <class M>
RTK_HD float weighted_coat_ior(const M& m) {
    return 1.0f + sat(m.coat_weight()) * (omax(m.coat_ior(), 1e-4f) - 1.0f);
}
RTK_HD float window1(float r) { return sat(r + sq(1.0f - r) * r / omax(1.0f - r * r, 1e-6f)); }
RTK_HD f3 window(f3 refl) {
    const f3 r = clamp_unit(refl);
    return v(window1(r.x), window1(r.y), window1(r.z));
}
template <class M>
RTK_HD f3 apply_film(const M& m, float cosine, f3 base, float inc_ior, f3 sub_ior, f3 gen_refl,
    bool gen_schlick) {
    const float w = sat(m.thin_film_weight());
    const float nm = omax(m.thin_film_thickness(), 0.0f) * 1000.0f;
    if (w <= 0.0f || nm <= 0.0f) return base;
    f3 film = film_airy(cosine, nm, inc_ior, m.thin_film_ior(), sub_ior, gen_refl, gen_schlick);
    if (m.geometry_thin_walled() != 0 && !gen_schlick) film = window(film);
    return lerp(base, film, w);
}
template <class M>
RTK_HD f3 dielectric_reflectance(const M& m, float cosine, bool entering) {
    if (!dispersion(m)) {
        const float base = dielectric_fresnel(m, cosine, entering);
        const float bior = modulated_ior(m);
        const float cior = weighted_coat_ior(m);
        const float inc = entering ? cior : bior;
        const float sb = entering ? bior : cior;
        return apply_film(m, cosine, vs(base), inc, vs(sb), vs(0.0f), false);
    }
    const f3 iors = dispersion_ior(m);
    const f3 eta = entering ? iors : recip(iors);
    f3 base = v(fresnel_dielectric(cosine, eta.x), fresnel_dielectric(cosine, eta.y),
        fresnel_dielectric(cosine, eta.z));
    if (m.geometry_thin_walled() != 0) base = window(base);
    const float cior = weighted_coat_ior(m);
    if (entering) return apply_film(m, cosine, base, cior, iors, vs(0.0f), false);
    f3 r = vs(0.0f);
    for (int c = 0; c < 3; ++c) {
        const f3 film = apply_film(m, cosine, base, comp(iors, c), vs(cior), vs(0.0f), false);
        set_comp(r, c, comp(film, c));
    }
    return r;
}
template <class M>
RTK_HD f3 metal_reflectance(const M& m, float cosine) {
    const f3 base = metal_fresnel(m, cosine);
    const float inc = weighted_coat_ior(m);
    const float ctt = film_cos_t(cosine, inc, m.thin_film_ior());
    const f3 at_film = metal_fresnel(m, ctt);
    const f3 normal = metal_fresnel(m, 1.0f);
    const f3 sb = v(inc * f0_to_ior(normal.x), inc * f0_to_ior(normal.y), inc * f0_to_ior(normal.z));
    return apply_film(m, cosine, base, inc, sb, at_film, true);
}
template <class M>
RTK_HD f3 dielectric_albedo_rgb(const M& m, float cosine, bool entering) {
    if ((m.thin_film_weight() <= 0.0f || m.thin_film_thickness() <= 0.0f) && !dispersion(m)) {
        const Rough r = spec_rough(m);
        const float alpha = avg_alpha(r);
        if (smooth_m(m)) return vs(dielectric_fresnel(m, cosine, entering));
        const float c0 = dielectric_fresnel(m, 1.0f, entering);
        if (c0 <= 1e-8f) return vs(0.0f);
        const float ss = ggx_albedo(cosine, alpha, c0, 1.0f);
        const float white = ggx_albedo(cosine, alpha, 1.0f, 1.0f);
        const float df = dielectric_fresnel(m, cosine, entering);
        const float cmp = 1.0f + df * (1.0f - white) / omax(white, 1e-4f);
        return vs(sat(ss * cmp));
    }
    const Rough r = spec_rough(m);
    const float alpha = avg_alpha(r);
    const f3 mirror = dielectric_reflectance(m, cosine, entering);
    if (smooth_m(m)) return mirror;
    const f3 c0 = dielectric_reflectance(m, 1.0f, entering);
    const f3 rough = v(ggx_albedo(cosine, alpha, c0.x, 1.0f), ggx_albedo(cosine, alpha, c0.y, 1.0f),
        ggx_albedo(cosine, alpha, c0.z, 1.0f));
    return lerp(mirror, rough, sqrtf(sat(alpha)));
}
template <class M>
RTK_HD f3 reflection_fresnel(const M& m, float mcos, float ocos, bool entering) {
    const float metal = sat(m.base_metalness());
    const float alpha = avg_alpha(spec_rough(m));
    const f3 df = dielectric_reflectance(m, mcos, entering);
    const f3 dsingle = mul(entering || m.geometry_thin_walled() != 0 ? clamp_unit(m.specular_color())
                                                                      : vs(1.0f), df);
    const f3 dielectric = mul(dsingle, energy_comp(ocos, alpha, df));
    const f3 msingle = metal_reflectance(m, mcos);
    const f3 metal_v = mul(msingle, energy_comp(ocos, alpha, msingle));
    return lerp(dielectric, metal_v, metal);
}

// ---- lobe selection
template <class M>
RTK_HD Probs lobe_probs(const M& m, float ocos, bool exterior) {
    Probs p{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    const float metal = sat(m.base_metalness());
    const float trans = sat(m.transmission_weight());
    const float bw = sat(m.base_weight());
    const float msw = sat(m.specular_weight());
    const float eta = modulated_ior(m);
    const float er = (eta - 1.0f) / (eta + 1.0f);
    const float df0 = er * er;
    float dsr = df0;
    float msr = msw * lum(scl(clamp_unit(m.base_color()), bw));
    if ((m.thin_film_weight() > 0.0f && m.thin_film_thickness() > 0.0f) || dispersion(m)) {
        dsr = lum(dielectric_reflectance(m, ocos, true));
        msr = lum(metal_reflectance(m, ocos));
    }
    const float opaque = (1.0f - metal) * (1.0f - trans) * sat(1.0f - dsr);
    const float ssw = sat(m.subsurface_weight());
    float diffuse = 0.0f, sss = 0.0f;
    if (ssw <= 0.0f) {
        diffuse = opaque * bw * lum(clamp_unit(m.base_color()));
    } else {
        const bool rw = ss_random_walk(m);
        const bool tw = ss_thin_walled(m);
        const float local_ss = rw || tw ? 0.0f : ssw;
        const f3 ldc = add(scl(clamp_unit(m.base_color()), bw * (1.0f - ssw)),
            scl(clamp_unit(m.subsurface_color()), local_ss));
        diffuse = opaque * lum(ldc);
        if (rw || tw) sss = opaque * ssw * lum(clamp_unit(m.subsurface_color()));
    }
    const float drs = (1.0f - metal) * dsr * lum(clamp_unit(m.specular_color()));
    const float mrs = metal * msr;
    const float ts = (1.0f - metal) * trans * sat(1.0f - dsr) * lum(surface_transmission_color(m));
    const float coat = exterior ? sat(m.coat_weight()) * coat_albedo(m, ocos) : 0.0f;
    const float fuzz = exterior ? sat(m.fuzz_weight()) * fuzz_albedo(ocos, m.fuzz_roughness())
                                      * lum(clamp_unit(m.fuzz_color()))
                                : 0.0f;
    const float scale = sat(1.0f - coat) * sat(1.0f - fuzz);
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

// ---- diffuse
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
template <class M>
RTK_HD f3 diffuse(const M& m, f3 wo, f3 wi) {
    if (wo.z <= 0.0f || wi.z <= 0.0f) return vs(0.0f);
    const float metal = sat(m.base_metalness());
    const float trans = sat(m.transmission_weight());
    const float ssw = sat(m.subsurface_weight());
    float mix = sat(m.base_weight());
    f3 color = clamp_unit(m.base_color());
    if (ssw > 0.0f) {
        const float bmix = sat(m.base_weight()) * (1.0f - ssw);
        const float smix = ss_random_walk(m) || ss_thin_walled(m) ? 0.0f : ssw;
        mix = bmix + smix;
        color = clamp_unit(odiv(add(scl(clamp_unit(m.base_color()), bmix),
                                    scl(clamp_unit(m.subsurface_color()), smix)),
            omax(mix, 1e-8f)));
    }
    const float coeff = mix * (1.0f - metal) * (1.0f - trans);
    if (coeff <= 0.0f) return vs(0.0f);
    const float k1 = 0.5f - 2.0f / (3.0f * kPiF);
    const float sigma = sat(m.base_diffuse_roughness());
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
template <class M>
RTK_HD f3 thin_ss_value(const M& m, f3 wo, f3 wi_same, bool transmission) {
    if (!ss_thin_walled(m)) return vs(0.0f);
    const MatThinDiffuse<M> d(m, sat(m.subsurface_weight()) * ss_thin_prob(m, transmission));
    return diffuse(d, wo, wi_same);
}

// ---- evaluate (evaluate_openpbr_core with make_openpbr_frame(n, 0))
template <class M>
RTK_HD void eval(const M& m, f3 n, f3 wo_world, f3 wi_world, f3 thr, f3* value, float* pdf) {
    *value = vs(0.0f);
    *pdf = 0.0f;
    const f3 won = onormalize(wo_world);
    const f3 win = onormalize(wi_world);
    if (lensq(won) <= 0.0f || lensq(win) <= 0.0f) return;
    Frame f = make_frame(n);
    const bool entering = m.geometry_thin_walled() != 0 || dot(f.n, won) >= 0.0f;
    f = face(f, won);
    const f3 wo = to_local(f, won);
    const f3 wi = to_local(f, win);
    if (wo.z <= 1e-7f || fabsf(wi.z) <= 1e-7f) return;
    const float opacity = sat(m.geometry_opacity());
    const bool exterior = entering;
    const Probs p = lobe_probs(m, wo.z, exterior);
    const Rough r = spec_rough(m);
    const bool is_smooth = smooth(r);
    f3 val = vs(0.0f);
    float pd = 0.0f;
    if (wi.z * wo.z > 0.0f) {
        const f3 att = sub(vs(1.0f), dielectric_albedo_rgb(m, wo.z, entering));
        val = add(diffuse(m, wo, wi), thin_ss_value(m, wo, wi, false));
        val = mul(val, clamp_unit(att));
        pd = p.diffuse * wi.z * kInvPi;
        if (ss_thin_walled(m)) pd += p.subsurface * ss_thin_prob(m, false) * wi.z * kInvPi;
        const f3 sm = add(wi, wo);
        if (!is_smooth && lensq(sm) > 1e-16f) {
            f3 wm = onormalize(sm);
            if (wm.z < 0.0f) wm = neg(wm);
            const float wodm = dot(wo, wm);
            if (wodm > 0.0f) {
                const float D = ggx_d(r, wm);
                const float G = ggx_g(r, wo, wi);
                const f3 F = reflection_fresnel(m, wodm, wo.z, entering);
                val = add(val, scl(F, D * G / omax(4.0f * wo.z * wi.z, 1e-8f)));
                pd += p.reflection * vndf_pdf(r, wo, wm) / omax(4.0f * fabsf(wodm), 1e-8f);
            }
        }
        val = mul(val, substrate_scale(m, wo.z, wi.z, exterior));
        if (exterior) {
            const f3 coat = scl(coat_value(m, wo, wi), fuzz_base_scale(m, wo.z, wi.z, true));
            val = add(val, coat);
            pd += p.coat * coat_pdf(m, wo, wi);
            if constexpr (M::kFull) {   // zero terms for the core
                val = add(val, fuzz_value(m, wo, wi));
                pd += p.fuzz * fuzz_pdf(m, wo, wi);
            }
        }
    } else if (ss_thin_walled(m) && p.subsurface > 0.0f) {
        const f3 wir = v(wi.x, wi.y, -wi.z);
        const f3 att = sub(vs(1.0f), dielectric_albedo_rgb(m, wo.z, true));
        val = thin_ss_value(m, wo, wir, true);
        val = mul(val, mul(clamp_unit(att), substrate_scale(m, wo.z, wi.z, true)));
        pd = p.subsurface * ss_thin_prob(m, true) * fabsf(wi.z) * kInvPi;
    } else if (M::kFull && !is_smooth && m.geometry_thin_walled() == 0 && p.transmission > 0.0f) {
        const float scale = (1.0f - sat(m.base_metalness())) * sat(m.transmission_weight());
        if (!dispersion(m)) {
            const float eta = modulated_ior(m);
            const float ep = entering ? eta : 1.0f / eta;
            const f3 sm = add(scl(wi, ep), wo);
            if (lensq(sm) > 1e-16f) {
                f3 wm = onormalize(sm);
                if (wm.z < 0.0f) wm = neg(wm);
                const float widm = dot(wi, wm);
                const float wodm = dot(wo, wm);
                if (widm * wi.z > 0.0f && wodm * wo.z > 0.0f) {
                    const f3 refl = dielectric_reflectance(m, wodm, entering);
                    const float hd = widm + wodm / ep;
                    const float d2 = hd * hd;
                    if (d2 > 1e-16f) {
                        const float D = ggx_d(r, wm);
                        const float G = ggx_g(r, wo, wi);
                        float t = D * G * fabsf(widm * wodm / (wi.z * wo.z * d2));
                        t /= ep * ep;
                        t *= scale;
                        val = scl(mul(surface_transmission_color(m), clamp_unit(sub(vs(1.0f), refl))), t);
                        const float deriv = fabsf(widm) / d2;
                        pd = p.transmission * vndf_pdf(r, wo, wm) * deriv;
                    }
                }
            }
        } else {
            const f3 iors = dispersion_ior(m);
            const f3 eps = entering ? iors : recip(iors);
            const f3 tc = surface_transmission_color(m);
            f3 cond = vs(0.0f);
            for (int c = 0; c < 3; ++c) {
                const float ep = comp(eps, c);
                const f3 sm = add(scl(wi, ep), wo);
                if (lensq(sm) <= 1e-16f) continue;
                f3 wm = onormalize(sm);
                if (wm.z < 0.0f) wm = neg(wm);
                const float widm = dot(wi, wm);
                const float wodm = dot(wo, wm);
                if (widm * wi.z <= 0.0f || wodm * wo.z <= 0.0f) continue;
                const float hd = widm + wodm / ep;
                const float d2 = hd * hd;
                if (d2 <= 1e-16f) continue;
                const float D = ggx_d(r, wm);
                const float G = ggx_g(r, wo, wi);
                float t = D * G * fabsf(widm * wodm / (wi.z * wo.z * d2));
                t *= scale / (ep * ep);
                const f3 refl = dielectric_reflectance(m, wodm, entering);
                set_comp(val, c, comp(tc, c) * sat(1.0f - comp(refl, c)) * t);
                const float deriv = fabsf(widm) / d2;
                set_comp(cond, c, vndf_pdf(r, wo, wm) * deriv);
            }
            pd = p.transmission * dot(cond, channel_probs(thr));
        }
        val = mul(val, substrate_scale(m, wo.z, wi.z, exterior));
    }
    *value = scl(val, opacity);
    *pdf = pd * opacity;
}

template <class M>
RTK_HD f3 emission(const M& m) {
    f3 r = scl(clamp_nonneg(m.emission_color()), omax(0.0f, m.emission_luminance()) * sat(m.geometry_opacity()));
    const float ct = 1.0f - sat(m.coat_weight()) * coat_albedo(m, 1.0f);
    r = mul(r, scl(coat_passage(m, 1.0f), sat(ct)));
    const float ft = 1.0f - sat(m.fuzz_weight()) * fuzz_albedo(1.0f, m.fuzz_roughness());
    return scl(r, sat(ft));
}

// ---- sample
struct Sample {
    f3 wi, value, weight;
    float pdf, discrete_pdf;
    int event;
    bool delta, valid;
};
RTK_HD Sample none() { return Sample{vs(0.0f), vs(0.0f), vs(0.0f), 0.0f, 0.0f, kPbrNone, false, false}; }

template <class M>
RTK_HD Sample finish(const M& m, f3 n, const Frame& f, f3 won, f3 wi, int event, f3 thr) {
    Sample s = none();
    s.wi = onormalize(to_world(f, wi));
    s.event = event;
    f3 val;
    float pdf;
    eval(m, n, won, s.wi, thr, &val, &pdf);
    if (pdf <= 1e-12f) return none();
    s.value = val;
    s.pdf = pdf;
    s.weight = scl(val, fabsf(dot(f.n, s.wi)) / pdf);
    s.valid = true;
    return s;
}
RTK_HD Sample delta_sample(const Frame& f, f3 wi, f3 coeff, float dp, int event, bool normalize) {
    Sample s = none();
    s.wi = normalize ? onormalize(to_world(f, wi)) : to_world(f, wi);
    s.discrete_pdf = dp;
    if (dp <= 0.0f) return none();
    s.weight = odiv(coeff, dp);
    s.event = event;
    s.delta = true;
    s.valid = true;
    return s;
}
RTK_HD f3 cos_hemisphere(float u1, float u2) {
    const float radius = sqrtf(sat(u1));
    float sp, cp;
    sincos_turn(sat(u2), &sp, &cp);
    return v(radius * cp, radius * sp, safe_sqrt(1.0f - radius * radius));
}

// openpbr_sample_subsurface_boundary over a boundary view
template <class M>
RTK_HD Sample sample_boundary(const M& m, f3 n, f3 wo_world, float ul, float u1, float u2, f3 thr) {
    const f3 won = onormalize(wo_world);
    if (lensq(won) <= 0.0f) return none();
    Frame f = make_frame(n);
    const bool entering = dot(f.n, won) >= 0.0f;
    f = face(f, won);
    const f3 wo = to_local(f, won);
    if (wo.z <= 1e-7f) return none();
    const Probs p = lobe_probs(m, wo.z, false);
    const float choice = oclamp(ul, 0.0f, 0.99999994f);
    const bool is_smooth = smooth_m(m);
    f3 wi;
    int event;
    if (choice < p.reflection) {
        if (is_smooth) {
            wi = v(-wo.x, -wo.y, wo.z);
            return delta_sample(f, wi, reflection_fresnel(m, wo.z, wo.z, entering), p.reflection,
                kPbrGlossyReflection, true);
        }
        const f3 wm = sample_vndf(spec_rough(m), wo, u1, u2);
        wi = reflect(wo, wm);
        if (wi.z <= 1e-7f) return none();
        event = kPbrGlossyReflection;
    } else {
        const float eta = modulated_ior(m);
        const float mep = entering ? eta : 1.0f / eta;
        float ep = 1.0f;
        f3 wm = v(0.0f, 0.0f, 1.0f);
        if (!is_smooth) wm = sample_vndf(spec_rough(m), wo, u1, u2);
        if (!refract(wo, wm, mep, &ep, &wi) || (!is_smooth && wi.z >= -1e-7f)) return none();
        if (is_smooth) {
            const f3 t = clamp_unit(sub(vs(1.0f), dielectric_reflectance(m, wo.z, entering)));
            return delta_sample(f, wi, scl(t, 1.0f / (ep * ep)), p.transmission, kPbrGlossyTransmission,
                true);
        }
        event = kPbrGlossyTransmission;
    }
    return finish(m, n, f, won, wi, event, thr);
}

template <class M>
RTK_HD Sample sample(const M& m, f3 n, f3 wo_world, float u_lobe, float u1, float u2, f3 thr) {
    const f3 won = onormalize(wo_world);
    if (lensq(won) <= 0.0f) return none();
    Frame f = make_frame(n);
    const bool entering = m.geometry_thin_walled() != 0 || dot(f.n, won) >= 0.0f;
    f = face(f, won);
    const f3 wo = to_local(f, won);
    if (wo.z <= 1e-7f) return none();
    const float opacity = sat(m.geometry_opacity());
    const float pass = 1.0f - opacity;
    u_lobe = sat(u_lobe);
    if (u_lobe < pass) {
        Sample s = none();
        s.wi = neg(won);
        s.weight = vs(1.0f);
        s.discrete_pdf = pass;
        s.event = kPbrPassthrough;
        s.delta = true;
        s.valid = true;
        return s;
    }
    if (opacity <= 0.0f) return none();
    const float choice = oclamp((u_lobe - pass) / opacity, 0.0f, 0.99999994f);
    if (!entering && ss_random_walk(m)) {
        // leaving the subsurface medium through its boundary
        Sample s = sample_boundary(MatBoundary<M>(m), n, won, choice, u1, u2, thr);
        if (!s.valid) return s;
        const f3 sc = substrate_scale(m, wo.z, dot(f.n, s.wi), false);
        s.value = mul(s.value, scl(sc, opacity));
        s.weight = mul(s.weight, sc);
        s.pdf *= opacity;
        s.discrete_pdf *= opacity;
        if (s.event == kPbrGlossyTransmission) s.event = kPbrSubsurfaceExit;
        return s;
    }
    const bool exterior = entering;
    const Probs p = lobe_probs(m, wo.z, exterior);
    if (p.diffuse + p.subsurface + p.reflection + p.transmission + p.coat + p.fuzz <= 0.0f) return none();
    const float fuzz_end = p.fuzz;
    const float coat_end = fuzz_end + p.coat;
    const float diffuse_end = coat_end + p.diffuse;
    const float ss_end = diffuse_end + p.subsurface;
    const float refl_end = ss_end + p.reflection;
    f3 wi;
    int event;
    if (M::kFull && choice < fuzz_end) {
        wi = sample_fuzz(m, wo, u1, u2);
        if (wi.z <= 1e-7f) return none();
        event = kPbrFuzz;
    } else if (choice < coat_end) {
        const Rough cr = coat_rough(m);
        if (smooth(cr)) {
            wi = v(-wo.x, -wo.y, wo.z);
            f3 coeff = vs(sat(m.coat_weight()) * coat_fresnel(m, wo.z) * opacity);
            coeff = scl(coeff, fuzz_base_scale(m, wo.z, wi.z, exterior));
            return delta_sample(f, wi, coeff, opacity * p.coat, kPbrCoat, true);
        }
        const f3 wm = sample_vndf(cr, wo, u1, u2);
        wi = reflect(wo, wm);
        if (wi.z <= 1e-7f) return none();
        event = kPbrCoat;
    } else if (choice < diffuse_end) {
        wi = cos_hemisphere(u1, u2);
        event = kPbrDiffuse;
    } else if (M::kFull && choice < ss_end) {
        if (ss_thin_walled(m)) {
            const float cc = oclamp((choice - diffuse_end) / omax(p.subsurface, 1e-8f), 0.0f, 0.99999994f);
            const float rp = ss_thin_prob(m, false);
            wi = cos_hemisphere(u1, u2);
            if (cc < rp) {
                event = kPbrDiffuse;
            } else {
                wi.z = -wi.z;
                event = kPbrThinTransmission;
            }
        } else {
            // entering the subsurface medium through its boundary
            const MatBoundary<M> b(m);
            const Probs bp = lobe_probs(b, wo.z, false);
            if (bp.transmission <= 0.0f || p.subsurface <= 0.0f) return none();
            const float cc = oclamp((choice - diffuse_end) / p.subsurface, 0.0f, 0.99999994f);
            const float bc = bp.reflection + cc * bp.transmission;
            Sample s = sample_boundary(b, n, won, bc, u1, u2, thr);
            if (!s.valid || s.event != kPbrGlossyTransmission) return none();
            const float ms = (1.0f - sat(m.base_metalness())) * (1.0f - sat(m.transmission_weight()))
                             * sat(m.subsurface_weight());
            const f3 ls = substrate_scale(m, wo.z, dot(f.n, s.wi), true);
            const f3 ws = scl(ls, ms * bp.transmission / omax(p.subsurface, 1e-8f));
            s.value = mul(s.value, scl(ls, ms * opacity));
            s.weight = mul(s.weight, ws);
            s.pdf *= opacity * p.subsurface / bp.transmission;
            s.discrete_pdf *= opacity * p.subsurface / bp.transmission;
            s.event = kPbrSubsurfaceEntry;
            return s;
        }
    } else if (choice < refl_end) {
        if (smooth_m(m)) {
            wi = v(-wo.x, -wo.y, wo.z);
            f3 coeff = scl(reflection_fresnel(m, wo.z, wo.z, entering), opacity);
            coeff = mul(coeff, substrate_scale(m, wo.z, wi.z, exterior));
            return delta_sample(f, wi, coeff, opacity * p.reflection, kPbrGlossyReflection, true);
        }
        const f3 wm = sample_vndf(spec_rough(m), wo, u1, u2);
        wi = reflect(wo, wm);
        if (wi.z <= 1e-7f) return none();
        event = kPbrGlossyReflection;
    } else if (!M::kFull) {
        // the core's transmission lobe has zero probability: the reference
        // yields an invalid sample on every branch past reflection_end
        return none();
    } else if (m.geometry_thin_walled() != 0) {
        wi = neg(wo);
        const f3 t = clamp_unit(sub(vs(1.0f), dielectric_reflectance(m, wo.z, true)));
        f3 coeff = scl(mul(surface_transmission_color(m), t),
            opacity * (1.0f - sat(m.base_metalness())) * sat(m.transmission_weight()));
        coeff = mul(coeff, substrate_scale(m, wo.z, wi.z, exterior));
        return delta_sample(f, wi, coeff, opacity * p.transmission, kPbrThinTransmission, false);
    } else if (smooth_m(m)) {
        if (dispersion(m)) {
            const float cu = oclamp((choice - refl_end) / omax(p.transmission, 1e-8f), 0.0f, 0.99999994f);
            const int ch = sample_channel(thr, cu);
            const f3 iors = dispersion_ior(m);
            const float mep = entering ? comp(iors, ch) : 1.0f / comp(iors, ch);
            float rep = 1.0f;
            if (!refract(wo, v(0.0f, 0.0f, 1.0f), mep, &rep, &wi)) return none();
            const f3 t = clamp_unit(sub(vs(1.0f), dielectric_reflectance(m, wo.z, entering)));
            const f3 tc = surface_transmission_color(m);
            f3 coeff = vs(0.0f);
            set_comp(coeff, ch, comp(tc, ch) * comp(t, ch) * opacity * (1.0f - sat(m.base_metalness()))
                                    * sat(m.transmission_weight()) / (rep * rep));
            coeff = mul(coeff, substrate_scale(m, wo.z, wi.z, exterior));
            const float cp = comp(channel_probs(thr), ch);
            return delta_sample(f, wi, coeff, opacity * p.transmission * cp, kPbrGlossyTransmission, true);
        }
        const float mep = entering ? modulated_ior(m) : 1.0f / modulated_ior(m);
        float rep = 1.0f;
        if (!refract(wo, v(0.0f, 0.0f, 1.0f), mep, &rep, &wi)) return none();
        const f3 t = clamp_unit(sub(vs(1.0f), dielectric_reflectance(m, wo.z, entering)));
        f3 coeff = scl(mul(surface_transmission_color(m), t),
            opacity * (1.0f - sat(m.base_metalness())) * sat(m.transmission_weight()) / (rep * rep));
        coeff = mul(coeff, substrate_scale(m, wo.z, wi.z, exterior));
        return delta_sample(f, wi, coeff, opacity * p.transmission, kPbrGlossyTransmission, true);
    } else {
        const f3 wm = sample_vndf(spec_rough(m), wo, u1, u2);
        float mep = entering ? modulated_ior(m) : 1.0f / modulated_ior(m);
        if (dispersion(m)) {
            const float cu = oclamp((choice - refl_end) / omax(p.transmission, 1e-8f), 0.0f, 0.99999994f);
            const int ch = sample_channel(thr, cu);
            const f3 iors = dispersion_ior(m);
            mep = entering ? comp(iors, ch) : 1.0f / comp(iors, ch);
        }
        float ep = 1.0f;
        if (!refract(wo, wm, mep, &ep, &wi) || wi.z >= -1e-7f) return none();
        event = kPbrGlossyTransmission;
    }
    return finish(m, n, f, won, wi, event, thr);
}

}  // namespace opbr

// ---- entry points over a material view (opbr::MatMem for the packed record)
template <class M>
RTK_HD f3 openpbr_emission_v(const M& m) { return opbr::emission(m); }
RTK_HD f3 openpbr_emission(const rt_openpbr_material& m) { return opbr::emission(opbr::MatMem{&m}); }

// evaluate with frame = make_openpbr_frame(n, 0); wo / wi world-space; thr is
// the path throughput of the transport context
template <class M>
RTK_HD void openpbr_eval_v(const M& m, f3 n, f3 wo_world, f3 wi_world, f3 thr, f3* value, float* pdf) {
    opbr::eval(m, n, wo_world, wi_world, thr, value, pdf);
}
RTK_HD void openpbr_eval(const rt_openpbr_material& m, f3 n, f3 wo_world, f3 wi_world, f3 thr,
    f3* value, float* pdf) {
    opbr::eval(opbr::MatMem{&m}, n, wo_world, wi_world, thr, value, pdf);
}

template <class M>
RTK_HD void openpbr_sample_v(const M& m, f3 n, f3 wo_world, float u_lobe, float u1, float u2,
    f3 thr, OpenPbrSampleResult* out) {
    const opbr::Sample s = opbr::sample(m, n, wo_world, u_lobe, u1, u2, thr);
    out->wi = s.wi;
    out->weight = s.weight;
    out->pdf = s.pdf;
    out->delta = s.delta;
    out->valid = s.valid;
    out->event = s.event;
}
RTK_HD void openpbr_sample(const rt_openpbr_material& m, f3 n, f3 wo_world, float u_lobe, float u1,
    float u2, f3 thr, OpenPbrSampleResult* out) {
    openpbr_sample_v(opbr::MatMem{&m}, n, wo_world, u_lobe, u1, u2, thr, out);
}

// the subsurface random walk: whether the material starts one, a free-flight
// segment toward a boundary `boundary` away, and the phase function's new
// direction (openpbr_core.h:407-413,428-519,576-590)
RTK_HD bool openpbr_random_walk(const rt_openpbr_material& m) {
    return opbr::ss_random_walk(opbr::MatMem{&m});
}
RTK_HD OpenPbrSegment openpbr_segment(const rt_openpbr_material& m, float boundary, float u, f3 thr) {
    return opbr::ss_segment(opbr::ss_medium(opbr::MatInside<opbr::MatMem>(opbr::MatMem{&m})), boundary, u, thr);
}
RTK_HD f3 openpbr_hg(const rt_openpbr_material& m, f3 incident, float u1, float u2) {
    return opbr::sample_hg(incident, opbr::ss_medium(opbr::MatInside<opbr::MatMem>(opbr::MatMem{&m})).g, u1, u2);
}

#undef RTK_HD

}  // namespace rtk
