// rt_openpbr.h -- device OpenPBR surface (constant inputs) for the shade stage.
#pragma once

#include "rt_internal.h"
#include "rt_math.h"

namespace rtk {

struct OpenPbrSampleResult {
    f3 wi, weight;
    float pdf;
    bool delta, valid;
    int event;
};

__device__ inline f3 openpbr_emission(const rt_openpbr_material& m) {
    (void)m;
    return mk3(0.0f, 0.0f, 0.0f);
}

__device__ inline void openpbr_eval(const rt_openpbr_material& m, f3 n, f3 wo, f3 wi, f3* value,
    float* pdf) {
    (void)m; (void)n; (void)wo; (void)wi;
    *value = mk3(0.0f, 0.0f, 0.0f);
    *pdf = 0.0f;
}

__device__ inline void openpbr_sample(const rt_openpbr_material& m, f3 n, f3 wo, float u_lobe,
    float u1, float u2, OpenPbrSampleResult* out) {
    (void)m; (void)n; (void)wo; (void)u_lobe; (void)u1; (void)u2;
    out->valid = false;
    out->pdf = 0.0f;
    out->delta = false;
    out->event = 0;
}

}  // namespace rtk
