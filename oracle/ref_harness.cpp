// ref_harness.cpp -- C entry points over the reference's dependency-free
// host/device headers, compiled in place from the reference tree into
// oracle/_ref/libref.so by `make ref`.  TEST INFRASTRUCTURE ONLY: used to pin
// the oracle's restatements (OpenPBR subset, light-sampling PDFs, ReSTIR
// reservoir math); never linked into the product.
//
// Wrapped headers: src/common/openpbr_core.h, src/common/light_sampling.h,
// src/common/restir_di.h, src/realtime/display_transfer.h.
#include "common/light_sampling.h"
#include "common/openpbr_core.h"
#include "common/restir_di.h"
#include "realtime/display_transfer.h"

#include <cstdint>
#include <cstring>

namespace {

// the oracle's rto_openpbr has OpenPbrCoreMaterial's field order (51 x 4 B)
rt::OpenPbrCoreMaterial to_ref(const float* p) {
    rt::OpenPbrCoreMaterial m{};
    int k = 0;
    auto f = [&]() { return p[k++]; };
    auto v = [&]() { rt::OpenPbrVec3 r{p[k], p[k + 1], p[k + 2]}; k += 3; return r; };
    m.base_weight = f(); m.base_color = v(); m.base_diffuse_roughness = f(); m.base_metalness = f();
    m.specular_weight = f(); m.specular_color = v(); m.specular_roughness = f();
    m.specular_ior = f(); m.specular_roughness_anisotropy = f();
    m.transmission_weight = f(); m.transmission_color = v(); m.transmission_depth = f();
    m.transmission_dispersion_scale = f(); m.transmission_dispersion_abbe_number = f();
    m.subsurface_weight = f(); m.subsurface_color = v(); m.subsurface_radius = f();
    m.subsurface_radius_scale = v(); m.subsurface_scatter_anisotropy = f();
    m.fuzz_weight = f(); m.fuzz_color = v(); m.fuzz_roughness = f();
    m.coat_weight = f(); m.coat_color = v(); m.coat_roughness = f();
    m.coat_roughness_anisotropy = f(); m.coat_ior = f(); m.coat_darkening = f();
    m.thin_film_weight = f(); m.thin_film_thickness = f(); m.thin_film_ior = f();
    m.emission_luminance = f(); m.emission_color = v();
    m.geometry_opacity = f();
    int32_t thin = 0;
    std::memcpy(&thin, p + k, 4);
    m.geometry_thin_walled = thin;
    return m;
}

rt::OpenPbrVec3 vec(const float* p) { return rt::OpenPbrVec3{p[0], p[1], p[2]}; }

}  // namespace

extern "C" {

int ref_openpbr_material_floats(void) { return static_cast<int>(sizeof(rt::OpenPbrCoreMaterial) / 4); }

// evaluate_openpbr_core with make_openpbr_frame(n, {}) as the integrator calls it
void ref_openpbr_evaluate(const float* material, const float* n, const float* wo, const float* wi,
    float* value_pdf) {
    const rt::OpenPbrCoreMaterial m = to_ref(material);
    const rt::OpenPbrFrame frame = rt::make_openpbr_frame(vec(n), rt::OpenPbrVec3{});
    const rt::OpenPbrEvaluation e = rt::evaluate_openpbr_core(m, frame, vec(wo), vec(wi));
    value_pdf[0] = e.value.x;
    value_pdf[1] = e.value.y;
    value_pdf[2] = e.value.z;
    value_pdf[3] = e.pdf;
}

// sample_openpbr_core; out = wi[3], weight[3], pdf, delta, valid, event
void ref_openpbr_sample(const float* material, const float* n, const float* wo, float u_lobe,
    float u1, float u2, float* out) {
    const rt::OpenPbrCoreMaterial m = to_ref(material);
    const rt::OpenPbrFrame frame = rt::make_openpbr_frame(vec(n), rt::OpenPbrVec3{});
    const rt::OpenPbrSample s = rt::sample_openpbr_core(m, frame, vec(wo), u_lobe, u1, u2);
    out[0] = s.wi.x; out[1] = s.wi.y; out[2] = s.wi.z;
    out[3] = s.weight.x; out[4] = s.weight.y; out[5] = s.weight.z;
    out[6] = s.pdf;
    out[7] = static_cast<float>(s.delta);
    out[8] = static_cast<float>(s.valid);
    out[9] = static_cast<float>(static_cast<int>(s.event));
}

// the same with the transport context's path throughput
void ref_openpbr_evaluate_ctx(const float* material, const float* n, const float* wo,
    const float* wi, const float* thr, float* value_pdf) {
    const rt::OpenPbrCoreMaterial m = to_ref(material);
    const rt::OpenPbrFrame frame = rt::make_openpbr_frame(vec(n), rt::OpenPbrVec3{});
    rt::OpenPbrTransportContext ctx{};
    ctx.path_throughput = vec(thr);
    const rt::OpenPbrEvaluation e = rt::evaluate_openpbr_core(m, frame, vec(wo), vec(wi), ctx);
    value_pdf[0] = e.value.x;
    value_pdf[1] = e.value.y;
    value_pdf[2] = e.value.z;
    value_pdf[3] = e.pdf;
}

void ref_openpbr_sample_ctx(const float* material, const float* n, const float* wo, float u_lobe,
    float u1, float u2, const float* thr, float* out) {
    const rt::OpenPbrCoreMaterial m = to_ref(material);
    const rt::OpenPbrFrame frame = rt::make_openpbr_frame(vec(n), rt::OpenPbrVec3{});
    rt::OpenPbrTransportContext ctx{};
    ctx.path_throughput = vec(thr);
    const rt::OpenPbrSample s = rt::sample_openpbr_core(m, frame, vec(wo), u_lobe, u1, u2, ctx);
    out[0] = s.wi.x; out[1] = s.wi.y; out[2] = s.wi.z;
    out[3] = s.weight.x; out[4] = s.weight.y; out[5] = s.weight.z;
    out[6] = s.pdf;
    out[7] = static_cast<float>(s.delta);
    out[8] = static_cast<float>(s.valid);
    out[9] = static_cast<float>(static_cast<int>(s.event));
}

// openpbr_subsurface_medium: extinction[3], albedo[3], anisotropy, active
void ref_openpbr_subsurface_medium(const float* material, float* out) {
    const rt::OpenPbrSubsurfaceMedium md = rt::openpbr_subsurface_medium(to_ref(material));
    out[0] = md.extinction.x; out[1] = md.extinction.y; out[2] = md.extinction.z;
    out[3] = md.single_scattering_albedo.x; out[4] = md.single_scattering_albedo.y;
    out[5] = md.single_scattering_albedo.z;
    out[6] = md.anisotropy;
    out[7] = static_cast<float>(md.active);
}

// openpbr_sample_subsurface_segment over the material's medium: weight[3], distance, scattered
void ref_openpbr_subsurface_segment(const float* material, float boundary, float u,
    const float* thr, float* out) {
    const rt::OpenPbrSubsurfaceMedium md = rt::openpbr_subsurface_medium(to_ref(material));
    rt::OpenPbrTransportContext ctx{};
    ctx.path_throughput = vec(thr);
    const rt::OpenPbrSubsurfaceSegment sg = rt::openpbr_sample_subsurface_segment(md, boundary, u, ctx);
    out[0] = sg.weight.x; out[1] = sg.weight.y; out[2] = sg.weight.z;
    out[3] = sg.distance;
    out[4] = static_cast<float>(sg.scattered);
}

void ref_openpbr_hg(const float* incident, float g, float u1, float u2, float* out) {
    const rt::OpenPbrVec3 d = rt::openpbr_sample_henyey_greenstein(vec(incident), g, u1, u2);
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
}

void ref_openpbr_emission(const float* material, float* out) {
    const rt::OpenPbrVec3 e = rt::emission_openpbr_core(to_ref(material));
    out[0] = e.x; out[1] = e.y; out[2] = e.z;
}

// light_sampling.h
float ref_light_uniform_cone_pdf(float c) { return rt::light_uniform_cone_pdf(c); }
float ref_light_area_to_solid_angle_pdf(float a, float d2, float c) {
    return rt::light_area_to_solid_angle_pdf(a, d2, c);
}
float ref_light_power_heuristic(float a, float b) { return rt::light_power_heuristic(a, b); }
int ref_sample_packed_light(const float* cdf, int count, float u) {
    rt::PackedLight lights[64];
    for (int i = 0; i < count && i < 64; ++i) lights[i].cdf = cdf[i];
    return rt::sample_packed_light(lights, count, u);
}
void ref_sample_uniform_triangle(float u0, float u1, float* w) {
    rt::sample_uniform_triangle(u0, u1, w[0], w[1], w[2]);
}

// restir_di.h reservoir: state = weight_sum, selected_target, estimator_weight,
// candidate_count, temporal_count, spatial_count, age, valid, selected light
void ref_restir_update(float* rs, int light, float target, float weight, int mult, float u,
    int* selected) {
    rt::RestirReservoir r{};
    r.weight_sum = rs[0]; r.selected_target = rs[1]; r.estimator_weight = rs[2];
    r.candidate_count = static_cast<int>(rs[3]); r.temporal_candidate_count = static_cast<int>(rs[4]);
    r.spatial_candidate_count = static_cast<int>(rs[5]); r.age = static_cast<int>(rs[6]);
    r.valid = static_cast<int>(rs[7]); r.selected.light_index = static_cast<int>(rs[8]);
    rt::RestirCandidate c{};
    c.light_index = light;
    *selected = rt::restir_update(r, c, target, weight, mult, u) ? 1 : 0;
    rs[0] = r.weight_sum; rs[1] = r.selected_target; rs[2] = r.estimator_weight;
    rs[3] = static_cast<float>(r.candidate_count); rs[4] = static_cast<float>(r.temporal_candidate_count);
    rs[5] = static_cast<float>(r.spatial_candidate_count); rs[6] = static_cast<float>(r.age);
    rs[7] = static_cast<float>(r.valid); rs[8] = static_cast<float>(r.selected.light_index);
}

void ref_restir_finalize(float* rs) {
    rt::RestirReservoir r{};
    r.weight_sum = rs[0]; r.selected_target = rs[1]; r.estimator_weight = rs[2];
    r.candidate_count = static_cast<int>(rs[3]); r.valid = static_cast<int>(rs[7]);
    r.selected.light_index = static_cast<int>(rs[8]);
    rt::restir_finalize(r);
    rs[2] = r.estimator_weight;
    rs[7] = static_cast<float>(r.valid);
}

float ref_linear_to_display(float v) { return rt::linear_to_display(v); }
uint8_t ref_linear_to_display_u8(float v) { return rt::linear_to_display_u8(v); }

}  // extern "C"
