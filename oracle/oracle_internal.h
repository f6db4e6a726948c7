/* oracle_internal.h -- declarations shared by the oracle's translation units.
 * TEST INFRASTRUCTURE ONLY (see rt_oracle.h). */
#ifndef RT_ORACLE_INTERNAL_H
#define RT_ORACLE_INTERNAL_H
#include "om.h"
#include "rt_oracle.h"

int om_closest(const rto_scene* s, const rto_bvh* bvh, v3 o, v3 d, float tmin, float tmax,
    float* t, int* prim, float* u, float* v);
int om_closest_counted(const rto_scene* s, const rto_bvh* bvh, v3 o, v3 d, float tmin, float tmax,
    float* t, int* prim, float* u, float* v, uint64_t* visits, uint64_t* tests);
int om_occluded(const rto_scene* s, const rto_bvh* bvh, v3 o, v3 d, float tmin, float tmax);

/* OpenPBR; thr is the path throughput of the transport context, which picks
 * the wavelength channel of dispersion and of the subsurface random walk */
v3 om_openpbr_emission(const rto_openpbr* m);
void om_openpbr_eval(const rto_openpbr* m, v3 n, v3 wo, v3 wi, v3 thr, v3* value, float* pdf);
void om_openpbr_sample(const rto_openpbr* m, v3 n, v3 wo, float u_lobe, float u1, float u2, v3 thr,
    rto_openpbr_sample* out);
int om_openpbr_random_walk(const rto_openpbr* m);
/* a free-flight segment of the subsurface random walk: out = weight[3],
 * distance, scattered */
void om_openpbr_segment(const rto_openpbr* m, float boundary, float u, v3 thr, float out[5]);
v3 om_openpbr_hg(const rto_openpbr* m, v3 incident, float u1, float u2);
#endif
