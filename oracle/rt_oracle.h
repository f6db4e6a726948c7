/*
 * rt_oracle.h -- CPU oracle for the path-tracing core.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against; it is never linked into, loaded by, or
 * called from the product library.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.
 *
 * It restates, in plain C11 on FP32 (FP64 for camera math, as the reference
 * does), the semantics of the reference's realtime GPU integrator
 *   src/realtime/gpu/programs.cu:1828-1953  (radiance_kernel)
 * on a counter-based RNG keyed on (pixel, sample, camera, bounce, dimension),
 * a Morton-code LBVH (Karras 2012) and brute-force closest hit.  Every
 * function cites the reference file:line it follows.
 *
 * Memory layouts of the structs below are identical to the C-ABI structs in
 * include/rt/rt_hip.h, so a test hands the same numpy buffers to both.
 *
 * Pinning (how each part is checked independently of the HIP path):
 *   RNG, sincos/cbrt  : closed-form/integer spec; KATs + math identities in
 *                       tests/test_oracle_math.py
 *   LBVH              : structural invariants + brute-force equivalence of
 *                       closest hits (tests/test_oracle_bvh.py)
 *   intersectors      : closed forms (tests/test_oracle_bvh.py)
 *   light sampling    : reference KATs of tests/test_light_sampling.cpp,
 *                       oracle/_ref light_sampling.h on random inputs,
 *                       sample pdf == hit pdf (tests/test_oracle_lights.py)
 *   camera models     : reference KATs of tests/test_camera_models.cpp and
 *                       the analytic forward models (tests/test_oracle_camera.py)
 *   BSDFs, integrator : reflection / Snell / Schlick laws, cosine-pdf
 *                       moments, furnaces and form-factor irradiance
 *                       closed forms, RR unbiasedness
 *                       (tests/test_oracle_integrator.py)
 *   OpenPBR subset    : oracle/_ref (reference header compiled here) +
 *                       reference KATs of tests/test_openpbr_core.cpp
 *   ReSTIR reservoir  : reference KATs of tests/test_restir_di.cpp
 */
#ifndef RT_ORACLE_H
#define RT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- scene ------------------------------------------------------------ */
typedef struct { float center[3]; float radius; int32_t material; } rto_sphere;
typedef struct { float origin[3]; float edge_u[3]; float edge_v[3]; int32_t material; } rto_quad;
typedef struct { float p0[3]; float p1[3]; float p2[3]; int32_t material; } rto_triangle;

enum { RTO_MAT_LAMBERTIAN = 0, RTO_MAT_METAL = 1, RTO_MAT_DIELECTRIC = 2,
       RTO_MAT_DIFFUSE_LIGHT = 3, RTO_MAT_ISOTROPIC = 4, RTO_MAT_OPENPBR = 5 };

typedef struct {
    int32_t type;
    float albedo[3];
    float emission[3];
    float fuzz;
    float ior;
    int32_t openpbr_index;
} rto_material;

/* Constant-input OpenPBR parameters (reference OpenPbrCoreMaterial,
 * src/common/openpbr_core.h:33-77, same field order, 51 floats/ints). */
typedef struct {
    float base_weight; float base_color[3]; float base_diffuse_roughness; float base_metalness;
    float specular_weight; float specular_color[3]; float specular_roughness; float specular_ior;
    float specular_roughness_anisotropy;
    float transmission_weight; float transmission_color[3]; float transmission_depth;
    float transmission_dispersion_scale; float transmission_dispersion_abbe_number;
    float subsurface_weight; float subsurface_color[3]; float subsurface_radius;
    float subsurface_radius_scale[3]; float subsurface_scatter_anisotropy;
    float fuzz_weight; float fuzz_color[3]; float fuzz_roughness;
    float coat_weight; float coat_color[3]; float coat_roughness; float coat_roughness_anisotropy;
    float coat_ior; float coat_darkening;
    float thin_film_weight; float thin_film_thickness; float thin_film_ior;
    float emission_luminance; float emission_color[3];
    float geometry_opacity; int32_t geometry_thin_walled;
} rto_openpbr;

/* PackedAnalyticLight (src/common/light_sampling.h:30-62), same field order */
enum { RTO_ALIGHT_SPHERE = 0, RTO_ALIGHT_DISK = 1, RTO_ALIGHT_RECT = 2, RTO_ALIGHT_CYLINDER = 3,
       RTO_ALIGHT_DISTANT = 4, RTO_ALIGHT_DOME = 5 };
typedef struct {
    int32_t type;
    float position[3];
    float basis_x[3];
    float basis_y[3];
    float basis_z[3];
    float radiance[3];
    float radius, width, height, length;
    float world_area;
    float cos_theta_max;
    float selection_pdf;
    float cdf;
    int32_t delta, treat_as_point, treat_as_line;
} rto_analytic_light;

/* PackedMedium (launch_params.h:201-214), same field order */
enum { RTO_MEDIUM_SPHERE = 0, RTO_MEDIUM_BOX = 1 };
typedef struct {
    float local_center_or_min[3];
    float radius;
    float local_max[3];
    float density;
    float rotation_row0[3];
    int32_t material_index;
    float rotation_row1[3];
    int32_t boundary_type;
    float rotation_row2[3];
    float pad0;
    float translation[3];
    float pad1;
} rto_medium;

enum { RTO_BG_CONSTANT = 0, RTO_BG_SKY_GRADIENT = 1 };

/* OpenPbrCompiledMaterial's texture bindings (openpbr_core.h:97-122):
 * texture index (-1 none) and, for colours, the source colour space
 * (OpenPbrSourceColorSpace: 0 raw, 1 linear sRGB, 2 sRGB texture) */
typedef struct {
    int32_t base_color, base_color_space;
    int32_t specular_color, specular_color_space;
    int32_t transmission_color, transmission_color_space;
    int32_t emission_color, emission_color_space;
    int32_t base_metalness, specular_roughness;
} rto_openpbr_textures;

/* rt_texture (PackedTexture, launch_params.h:225-237) */
typedef struct {
    int32_t type, even_texture, odd_texture;
    int32_t image_offset, image_width, image_height;
    int32_t u_address_mode, v_address_mode, filter_type;
    float scale;
    float color[3];
} rto_texture;

typedef struct {
    int32_t sphere_count, quad_count, triangle_count, material_count;
    int32_t openpbr_count, analytic_light_count;
    int32_t background_mode;
    float background[3];     /* constant colour, or gradient colour at d.y = -1 */
    float background_top[3]; /* gradient colour at d.y = +1 */
    const rto_sphere* spheres;
    const rto_quad* quads;
    const rto_triangle* triangles;
    const rto_material* materials;
    const rto_openpbr* openpbr;
    const rto_analytic_light* analytic_lights;
    int32_t texture_count, image_texel_count;
    const rto_texture* textures;
    const float* image_texels;
    const int32_t* material_albedo_texture;
    const float* triangle_uvs;
    const float* triangle_normals;
    int32_t medium_count;
    int32_t pad_media;
    const rto_medium* media;
    const int32_t* material_emission_texture;       /* MaterialSample::emission_texture, or NULL */
    const rto_openpbr_textures* openpbr_textures;   /* OpenPbrCompiledMaterial bindings, or NULL */
} rto_scene;

/* ---- cameras / profile ------------------------------------------------ */
enum { RTO_CAM_PINHOLE32 = 0, RTO_CAM_EQUI62_LUT1D = 1 };
typedef struct {
    int32_t enabled, width, height, model;
    double origin[3], basis_x[3], basis_y[3], basis_z[3];
    double fx, fy, cx, cy;
    double k1, k2, k3, p1, p2;
    double radial[6], tangential[2];
    double lut_step;
    double lut[1024];
} rto_camera;

typedef struct {
    int32_t samples_per_pixel, max_bounces, enable_denoise, rr_start_bounce;
    double accumulation_reset_rotation_deg, accumulation_reset_translation;
    int32_t enable_restir_di, restir_initial_candidates, restir_temporal_reuse,
        restir_max_history_age, restir_max_temporal_candidates, restir_spatial_neighbors,
        restir_max_spatial_candidates, restir_bias_correction, restir_min_analytic_lights;
    int32_t pixel_jitter;          /* extension: jitter primary rays inside the pixel */
    int32_t restir_spatial_radius; /* extension: spatial reuse radius in pixels */
    int32_t pad0;
} rto_profile;

/* ---- BVH (binary LBVH, child boxes stored in the parent) -------------- */
/* 64-byte node: child 0 box, child 1 box, two child references.
 * A child reference with bit 31 set is a leaf: low 31 bits = position in
 * the Morton-sorted primitive order.  0xFFFFFFFF = empty child. */
typedef struct {
    float lo0[3]; float hi0[3];
    float lo1[3]; float hi1[3];
    uint32_t child[2];
    uint32_t parent;
    uint32_t pad;
} rto_bvh_node;

/* 4-wide quantized node collapsed from the binary LBVH (64 B): origin and
 * per-axis quantum exponents (biased by 127) + child count in meta, child
 * references (leaf bit | sorted leaf slot, wide node index, or empty), and
 * per-axis child bounds in quanta (byte k of each word = child k) */
typedef struct {
    float origin[3];
    uint32_t meta;            /* ex | ey << 8 | ez << 16 | count << 24 */
    uint32_t child[4];
    uint32_t qlo[3];
    uint32_t qhi[3];
    uint32_t pad[2];
} rto_wide_node;

typedef struct {
    int32_t prim_count;       /* spheres + quads + triangles */
    int32_t node_count;       /* max(prim_count - 1, 1) if prim_count > 0 */
    rto_bvh_node* nodes;
    uint32_t* sorted_prims;   /* leaf position -> global primitive id */
    uint32_t* morton;         /* sorted Morton codes */
    int32_t max_depth;        /* internal levels on the deepest root-leaf path */
    int32_t wide_count;       /* 4-wide nodes */
    rto_wide_node* wide;
    int32_t wide_levels;      /* levels of the 4-wide tree */
    int32_t pad;
} rto_bvh;

typedef struct { float o[3]; float tmin; float d[3]; float tmax; } rto_ray;
typedef struct { float t; int32_t prim; float u; float v; } rto_hit;

/* ---- API -------------------------------------------------------------- */
int rto_version(void);

/* RNG: counter hash (Jarzynski & Olano 2020, pcg4d) */
void rto_pcg4d(uint32_t v[4]);
void rto_rng4(uint32_t pixel, uint32_t sample, uint32_t camera, uint32_t bounce, uint32_t group,
    uint32_t stream, float out[4]);
void rto_sincos_turn(float u, float* s, float* c);
float rto_cbrt01(float x);

/* scene generators (seeded PCG32) */
uint32_t rto_pcg32_next(uint64_t* state, uint64_t inc);
int rto_build_lbvh(const rto_scene* scene, rto_bvh* out);   /* allocates */
/* new primitive positions over the topology and leaf order of `bvh` */
int rto_refit_lbvh(const rto_scene* scene, rto_bvh* bvh);
void rto_free_lbvh(rto_bvh* bvh);
int rto_bvh_validate(const rto_scene* scene, const rto_bvh* bvh);
uint32_t rto_expand_bits10(uint32_t v);

/* closest hit / any hit; counters (node visits, prim tests) may be NULL */
void rto_trace_closest(const rto_scene* scene, const rto_bvh* bvh, const rto_ray* rays, int n,
    rto_hit* hits, uint32_t* node_visits, uint32_t* prim_tests, int threads);
void rto_trace_any(const rto_scene* scene, const rto_bvh* bvh, const rto_ray* rays, int n,
    int32_t* occluded, uint32_t* node_visits, uint32_t* prim_tests, int threads);
void rto_brute_closest(const rto_scene* scene, const rto_ray* rays, int n, rto_hit* hits);

/* cameras */
void rto_unproject(const rto_camera* cam, double px, double py, float dir_cam[3]);
void rto_camera_ray(const rto_camera* cam, double px, double py, float origin[3], float dir[3]);
void rto_project(const rto_camera* cam, const float dir_cam[3], float pixel[2]);
/* the camera-model debug image: width * height * 4 bytes, RGBA8 encoded
 * world directions of the pixel centres' rays (programs.cu:1807-1826) */
void rto_direction_debug(const rto_camera* cam, uint8_t* rgba);
/* the fixed-sequence natural logarithm the GPU uses for free-flight
 * distances (test hook) */
float rto_log_fixed(float x);
void rto_make_equi_lut(rto_camera* cam);
/* the orbiting rig of the instanced scene (generator id 2 or 3) at frame `frame` of `frames` */
int rto_generate_rig(int config, uint64_t seed, int frame, int frames, rto_camera* cams, int* count);

/* ---- learned denoiser (oracle_nn.c) */
typedef struct { int32_t count; int32_t offset[17]; int32_t width[16]; } rto_images;
int rto_nn_param_count(void);
float rto_bf16_round(float x);
void rto_nn_features(int P, const float* beauty, const float* normal, const float* albedo,
    const float* depth, float* feat);
double rto_nn_loss_grad(const rto_images* im, const float* params, const float* feat,
    const float* beauty, const float* target, float* out4, float* grad, int bf16);

/* ---- radiance cache (oracle_cache.c) */
typedef struct rto_cache_grid {
    int32_t levels, log2_table;
    float res[16];
    int32_t dense[16];
    float lo[3], inv_extent[3];
} rto_cache_grid;
int rto_cache_mlp_param_count(void);
/* table indices and trilinear weights of level l around p in [0,1]^3 */
void rto_cache_corners(const rto_cache_grid* g, int l, const float p[3], uint32_t idx[8], float w[8]);
/* the 64 network inputs of one 16-float query */
void rto_cache_features(const rto_cache_grid* g, const float* table, const float* in16, int bf16,
    float x[64]);
/* predictions (n x 3, double) for n 16-float queries; params = MLP then grid */
int rto_cache_predict(const rto_cache_grid* g, const float* params, int n, const float* inputs,
    int bf16, double* out);
/* relative-L2 loss of n samples (targets n x 4) and, when grad is set, its
 * gradient (MLP then grid, double) */
double rto_cache_loss_grad(const rto_cache_grid* g, const float* params, int n, const float* inputs,
    const float* targets, int bf16, double* grad);

/* ---- per-frame post passes (oracle_post.c) */
/* temporal reprojection + accumulation of one camera image, in place */
void rto_resolve(const rto_camera* cam, const rto_camera* prev_cam, int history_length,
    int restir_enabled, float* beauty, float* normal, float* depth, const float* hbeauty,
    const float* hnormal, const float* hdepth, float* flow, float* trust);
/* one edge-aware a-trous pass (spacing `step`) / the full filter */
void rto_denoise_pass(int w, int h, int step, const float* in, const float* normal,
    const float* albedo, const float* depth, float* out);
void rto_denoise(int w, int h, int iterations, const float* beauty, const float* normal,
    const float* albedo, const float* depth, float* out, float* scratch);
float rto_noise_sin(float x);
float rto_turbulence(const float p[3]);
void rto_eval_texture(const rto_scene* s, int index, float u, float v, const float p[3], float out[3]);
void rto_surface_uv(const rto_scene* s, const float o[3], const float d[3], int prim_type,
    int prim_index, const float pos[3], float* u, float* v);
uint8_t rto_display_u8(float v);
void rto_tone_map_rgba8(const float* rgba, int64_t n, uint8_t* out);
void rto_primary_rays(const rto_camera* cam, int cam_index, int sample, uint32_t stream,
    int jitter, rto_ray* rays);

/* light distribution (reference packed_scene_preparation.cpp:265-302) */
typedef struct { int32_t type; int32_t prim; float selection_pdf; float cdf; } rto_light;
int rto_build_lights(const rto_scene* scene, rto_light* out, int capacity);

/* full render of one camera; outputs are W*H*4 floats (beauty, normal,
 * albedo) and W*H floats (depth).  Any output may be NULL. */
/* node_visits / prim_tests: child boxes tested and primitives tested by the
 * path-extension (closest-hit) traversals, the counts the GPU's counting
 * frame reports */
typedef struct {
    uint64_t paths; uint64_t rays_closest; uint64_t rays_shadow;
    uint64_t node_visits; uint64_t prim_tests;
} rto_render_stats;
int rto_render_camera(const rto_scene* scene, const rto_bvh* bvh, const rto_camera* cam,
    int cam_index, const rto_profile* profile, uint32_t stream, float* beauty, float* normal,
    float* albedo, float* depth, int threads, int row_begin, int row_end,
    rto_render_stats* stats);
/* ReSTIR DI reservoir, field order of RestirReservoir (src/common/restir_di.h:36-48), 84 B */
typedef struct {
    int32_t light_index; float sample_u0, sample_u1;
    float position[3]; float normal[3];
    int32_t material_type, primitive_type, primitive_index;
    float weight_sum, selected_target, estimator_weight;
    int32_t candidate_count, temporal_candidate_count, spatial_candidate_count,
        bias_correction_source_count, age, valid;
} rto_reservoir;

/* per-camera frame history for ReSTIR reuse; any pointer may be NULL */
typedef struct {
    const rto_camera* prev_cam;     /* camera of the previous frame (NULL: no history) */
    const rto_reservoir* history;   /* previous frame's reservoirs, W*H */
    rto_reservoir* out;             /* this frame's reservoirs, W*H (last sample wins) */
    int32_t history_length;
    int32_t pad;
} rto_restir_io;

int rto_render_camera_ex(const rto_scene* scene, const rto_bvh* bvh, const rto_camera* cam,
    int cam_index, const rto_profile* profile, uint32_t stream, const rto_restir_io* restir,
    float* beauty, float* normal, float* albedo, float* depth, int threads, int row_begin,
    int row_end, rto_render_stats* stats);
/* the analytic light compiler's derived fields and selection distribution
 * (src/scene/analytic_light_compiler.cpp:15-39,114-128,178-240): world area
 * per type, delta distant lights, FP64 selection weights normalized into
 * selection_pdf / cdf.  Returns 0, or -1 for an unknown type or a sphere /
 * cylinder whose basis is not a similarity transform. */
int rto_analytic_lights_compile(rto_analytic_light* lights, int count);

/* reservoir math (src/common/restir_di.h:67-199) for known-answer tests */
int rto_restir_update(rto_reservoir* r, int light, float u0, float u1, float target, float weight,
    int multiplicity, float u);
void rto_restir_finalize(rto_reservoir* r);
void rto_restir_finalize_basic(rto_reservoir* r, float selected_source_target,
    float source_target_sum, int source_count);
int rto_restir_merge(rto_reservoir* cur, const rto_reservoir* prev, float target, int max_count,
    float u, int temporal, int* selected);

/* per-path radiance of one (pixel, sample): for tests of single paths */
void rto_trace_path(const rto_scene* scene, const rto_bvh* bvh, const rto_camera* cam,
    int cam_index, const rto_profile* profile, uint32_t stream, int x, int y, int sample,
    float radiance[3]);

/* pieces of the integrator, exposed for known-answer and closed-form tests */
float rto_power_heuristic(float a, float b);                  /* light_sampling.h:77-84 */
float rto_cone_pdf(float cos_max);                            /* light_sampling.h:64-67 */
float rto_area_pdf(float area, float dsq, float abs_cos);     /* light_sampling.h:69-75 */
float rto_uniform_sphere_pdf(void);                           /* light_sampling.h:60-62 */
int rto_pick_cdf(const float* cdf, int n, float u);           /* light_sampling.h:87-101 */
void rto_triangle_barycentrics(float u0, float u1, float w[3]); /* light_sampling.h:118-125 */
/* make_hit's base colour, emission and texture coordinates for the ray
 * o + t d on global primitive prim (test hook) */
void rto_hit_material(const rto_scene* s, const float o[3], const float d[3], float t, int prim,
    float out[8]);
int rto_sample_light_at(const rto_scene* s, const rto_bvh* bvh, int prim, const float pos[3],
    const float gn[3], const float d[3], const float r[4], float out[8]);
float rto_light_pdf_for_hit(const rto_scene* s, const rto_bvh* bvh, int prim, int prev_mat,
    const float pos[3], const float sn[3], const float dir[3]);
int rto_scatter(const rto_scene* s, int prim, const float pos[3], const float gn[3],
    const float d[3], const float r[4], float out[8]);

/* OpenPBR (src/common/openpbr_core.h); the _ctx forms take the path
 * throughput of the transport context, the others use (1, 1, 1) */
typedef struct { float value[3]; float pdf; } rto_openpbr_eval;
typedef struct { float wi[3]; float weight[3]; float pdf; int32_t delta; int32_t valid; int32_t event; } rto_openpbr_sample;
void rto_openpbr_evaluate(const rto_openpbr* m, const float n[3], const float wo[3],
    const float wi[3], rto_openpbr_eval* out);
void rto_openpbr_emission(const rto_openpbr* m, float out[3]);
void rto_openpbr_sample_fn(const rto_openpbr* m, const float n[3], const float wo[3],
    float u_lobe, float u1, float u2, rto_openpbr_sample* out);
void rto_openpbr_evaluate_ctx(const rto_openpbr* m, const float n[3], const float wo[3],
    const float wi[3], const float thr[3], rto_openpbr_eval* out);
void rto_openpbr_sample_ctx(const rto_openpbr* m, const float n[3], const float wo[3],
    float u_lobe, float u1, float u2, const float thr[3], rto_openpbr_sample* out);
/* openpbr_subsurface_medium: extinction[3], albedo[3], anisotropy, active */
void rto_openpbr_subsurface_medium(const rto_openpbr* m, float out[8]);
/* openpbr_sample_subsurface_segment: weight[3], distance, scattered */
void rto_openpbr_subsurface_segment(const rto_openpbr* m, float boundary, float u,
    const float thr[3], float out[5]);
/* openpbr_sample_henyey_greenstein */
void rto_openpbr_hg(const float incident[3], float g, float u1, float u2, float out[3]);

#ifdef __cplusplus
}
#endif
#endif
