/* oracle_openpbr.c -- OpenPBR subset (src/common/openpbr_core.h).
 * TEST INFRASTRUCTURE ONLY (see rt_oracle.h). */
#include "oracle_internal.h"

v3 om_openpbr_emission(const rto_openpbr* m) {
    (void)m;
    return mk3(0.0f, 0.0f, 0.0f);
}
void om_openpbr_eval(const rto_openpbr* m, v3 n, v3 wo, v3 wi, v3* value, float* pdf) {
    (void)m; (void)n; (void)wo; (void)wi;
    *value = mk3(0.0f, 0.0f, 0.0f);
    *pdf = 0.0f;
}
void om_openpbr_sample(const rto_openpbr* m, v3 n, v3 wo, float u_lobe, float u1, float u2,
    rto_openpbr_sample* out) {
    (void)m; (void)n; (void)wo; (void)u_lobe; (void)u1; (void)u2;
    memset(out, 0, sizeof(*out));
}
