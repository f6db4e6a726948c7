/*
 * rt_hip.h -- C ABI of the MI355X (gfx950) path-tracing core.
 *
 * Plain C: POD structs, opaque handles, explicit sizes, rt_status return
 * codes and a thread-local rt_last_error().  No HIP or C++ types cross this
 * boundary.  Each entry point names the reference interface it replaces
 * (paths relative to the reference repository root).
 *
 * Threading: handles are independent; calls on one handle must not overlap.
 * Ownership: handles own their device memory; callers own every host array.
 * A scene keeps its device alive and a renderer keeps its scene alive, so the
 * *_destroy calls may come in any order: each object is freed when the last
 * handle using it is destroyed.
 */
#ifndef RT_HIP_H
#define RT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RT_ABI_VERSION 3

typedef enum rt_status {
    RT_OK = 0,
    RT_ERROR_INVALID_ARGUMENT = 1, /* std::invalid_argument in the reference */
    RT_ERROR_HIP = 2,              /* "CUDA runtime failure at ..." in the reference */
    RT_ERROR_NOT_READY = 3,        /* e.g. render before prepare_scene() */
    RT_ERROR_OUT_OF_MEMORY = 4,
    RT_ERROR_UNSUPPORTED = 5
} rt_status;

/* ---------------------------------------------------------------- scene
 * Flattened world-space primitives, as PackedScene after
 * adapt_scene_ir_v2_to_realtime (src/realtime/scene_description.h:130-146,
 * src/scene/realtime_scene_adapter.cpp:371-487).  Global primitive ids are
 * spheres, then quads, then triangles. */
typedef struct rt_sphere { float center[3]; float radius; int32_t material; } rt_sphere;
typedef struct rt_quad { float origin[3]; float edge_u[3]; float edge_v[3]; int32_t material; } rt_quad;
typedef struct rt_triangle { float p0[3]; float p1[3]; float p2[3]; int32_t material; } rt_triangle;

/* material types of MaterialSample (src/realtime/gpu/launch_params.h:216-223) */
enum { RT_MAT_LAMBERTIAN = 0, RT_MAT_METAL = 1, RT_MAT_DIELECTRIC = 2, RT_MAT_DIFFUSE_LIGHT = 3,
       RT_MAT_ISOTROPIC = 4, RT_MAT_OPENPBR = 5 };

typedef struct rt_material {
    int32_t type;
    float albedo[3];
    float emission[3];
    float fuzz;
    float ior;
    int32_t openpbr_index;
} rt_material;

/* PackedTexture (src/realtime/gpu/launch_params.h:225-237), same field
 * order.  Types: 0 constant colour; 1 checker over the world position
 * (floor(p / scale) parity picks even_texture or odd_texture, chains resolve
 * at most 8 deep); 2 image (image_width x image_height texels from
 * image_offset in rt_scene_desc.image_texels; address modes 0 border (black
 * outside [0, 1]), 1 clamp, 2 repeat, 3 mirror; filter 0 nearest,
 * 1 bilinear); 3 turbulence noise.  An index outside the table or a chain
 * deeper than 8 evaluates to magenta (1, 0, 1), as in evaluate_texture
 * (src/realtime/gpu/programs.cu:268-302). */
enum { RT_TEX_CONSTANT = 0, RT_TEX_CHECKER = 1, RT_TEX_IMAGE = 2, RT_TEX_NOISE = 3 };
typedef struct rt_texture {
    int32_t type, even_texture, odd_texture;
    int32_t image_offset, image_width, image_height;
    int32_t u_address_mode, v_address_mode, filter_type;
    float scale;
    float color[3];
} rt_texture;

/* OpenPbrCoreMaterial, src/common/openpbr_core.h:33-77 (same field order) */
typedef struct rt_openpbr_material {
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
} rt_openpbr_material;

/* The texture bindings of OpenPbrCompiledMaterial (src/common/openpbr_core.h:
 * 97-122), one record per OpenPBR material.  Each field is a texture index
 * (-1: the material's constant input).  A colour binding carries the colour
 * space of its texture: RT_COLOR_SRGB_TEXTURE values are decoded with the
 * sRGB transfer function, the others are taken as linear
 * (openpbr_source_to_linear, openpbr_core.h:322-333).  base_color and
 * emission_color are evaluated at every hit (populate_material,
 * programs.cu:995-1008) and at light samples; specular_color,
 * transmission_color, base_metalness and specular_roughness when the surface
 * scatters (evaluate_openpbr_scattering_material, programs.cu:944-959).  A
 * scalar binding reads the texture's first channel. */
enum { RT_COLOR_RAW = 0, RT_COLOR_LINEAR_SRGB = 1, RT_COLOR_SRGB_TEXTURE = 2 };
typedef struct rt_openpbr_textures {
    int32_t base_color, base_color_space;
    int32_t specular_color, specular_color_space;
    int32_t transmission_color, transmission_color_space;
    int32_t emission_color, emission_color_space;
    int32_t base_metalness, specular_roughness;
} rt_openpbr_textures;

/* PackedAnalyticLight (src/common/light_sampling.h:30-62), same field order.
 * basis_x / basis_y / basis_z are the columns of the light's local-to-world
 * linear map.  Sphere: radius scaled by |basis_x|; disk: radius in the
 * basis_x / basis_y plane; rect: width x height along basis_x / basis_y;
 * cylinder: radius across basis_x / basis_y, length along basis_z; all emit
 * on the side facing -(basis_x x basis_y) (disk, rect) or outward (sphere,
 * cylinder).  Distant: a cone of half-angle acos(cos_theta_max) around
 * basis_z, delta != 0 for a single direction; dome: every direction.
 * world_area, selection_pdf and cdf are the compiled values the kernels
 * read as given (rt_analytic_lights_compile computes them); treat_as_point
 * makes a sphere a point light; treat_as_line drops a finite light from
 * emission MIS. */
enum { RT_ALIGHT_SPHERE = 0, RT_ALIGHT_DISK = 1, RT_ALIGHT_RECT = 2, RT_ALIGHT_CYLINDER = 3,
       RT_ALIGHT_DISTANT = 4, RT_ALIGHT_DOME = 5 };
typedef struct rt_analytic_light {
    int32_t type;
    float position[3];
    float basis_x[3];
    float basis_y[3];
    float basis_z[3];
    float radiance[3];
    float radius;
    float width;
    float height;
    float length;
    float world_area;
    float cos_theta_max;
    float selection_pdf;
    float cdf;
    int32_t delta;
    int32_t treat_as_point;
    int32_t treat_as_line;
} rt_analytic_light;

/* PackedMedium (src/realtime/gpu/launch_params.h:201-214), same field order:
 * a homogeneous medium of the given density inside a sphere (boundary_type
 * 0: local_center_or_min + radius) or a box (1: local_center_or_min /
 * local_max), in the medium's frame: local = R (p - translation) with R's
 * rows rotation_row0..2.  A closest-hit ray that crosses it scatters at an
 * exponentially distributed free-flight distance, with the medium's material
 * (normally RT_MAT_ISOTROPIC); shadow rays pass through media, as in the
 * reference (is_occluded, programs.cu:1759-1762).  At most 60 media. */
enum { RT_MEDIUM_SPHERE = 0, RT_MEDIUM_BOX = 1 };
typedef struct rt_medium {
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
} rt_medium;

enum { RT_BG_CONSTANT = 0, RT_BG_SKY_GRADIENT = 1 };

typedef struct rt_scene_desc {
    int32_t sphere_count, quad_count, triangle_count, material_count;
    int32_t openpbr_count, analytic_light_count;
    int32_t background_mode;
    float background[3];     /* constant colour, or gradient colour at d.y = -1 */
    float background_top[3]; /* gradient colour at d.y = +1 */
    const rt_sphere* spheres;
    const rt_quad* quads;
    const rt_triangle* triangles;
    const rt_material* materials;
    const rt_openpbr_material* openpbr;
    const rt_analytic_light* analytic_lights;
    /* textures.  material_albedo_texture[i] >= 0 replaces material i's
     * constant albedo by that texture evaluated at the hit
     * (MaterialSample::albedo_texture, launch_params.h:216-218); NULL or -1
     * keeps the constant.  Texture coordinates: sphere_uv of the outward
     * normal, the quad's edge coordinates, the triangle's barycentrics
     * (programs.cu:1447,1485,1528). */
    int32_t texture_count, image_texel_count;
    const rt_texture* textures;
    const float* image_texels;               /* 3 floats (linear RGB) per texel */
    const int32_t* material_albedo_texture;  /* material_count entries, or NULL */
    /* per-triangle texture coordinates uv0, uv1, uv2 (6 floats per
     * triangle; PackedTriangle::uv0..uv2 with has_texcoords), or NULL.  A
     * triangle whose first float is NaN has none and uses its barycentrics. */
    const float* triangle_uvs;
    /* per-triangle vertex normals n0, n1, n2 (9 floats per triangle;
     * PackedTriangle::n0..n2 with has_vertex_normals), or NULL; a NaN first
     * float means none.  They set the shading normal (programs.cu:1535-1545). */
    const float* triangle_normals;
    /* homogeneous media (PackedScene::media), or none */
    int32_t medium_count;
    int32_t pad_media;
    const rt_medium* media;
    /* emission textures.  material_emission_texture[i] >= 0 replaces
     * material i's constant emission by that texture, evaluated at the hit
     * and at the sampled point of a light (MaterialSample::emission_texture,
     * programs.cu:980-981,1225-1226); the light-selection power takes the
     * texture's average (packed_scene_preparation.cpp:195-224).  Only diffuse
     * lights emit.  NULL or -1 keeps the constant. */
    const int32_t* material_emission_texture;  /* material_count entries, or NULL */
    /* OpenPBR texture bindings, openpbr_count records, or NULL (none bound) */
    const rt_openpbr_textures* openpbr_textures;
} rt_scene_desc;

/* -------------------------------------------------------------- cameras
 * DeviceActiveCamera (src/realtime/gpu/launch_params.h:63-98): basis_* are
 * the columns of T_rc's rotation (renderer <- camera, OpenCV axes). */
enum { RT_CAM_PINHOLE32 = 0, RT_CAM_EQUI62_LUT1D = 1 };
typedef struct rt_camera {
    int32_t enabled, width, height, model;
    double origin[3], basis_x[3], basis_y[3], basis_z[3];
    double fx, fy, cx, cy;
    double k1, k2, k3, p1, p2;
    double radial[6], tangential[2];
    double lut_step;
    double lut[1024];
} rt_camera;

/* fills lut_step / lut of an equi62_lut1d camera from its size, intrinsics
 * and radial coefficients (make_equi62_lut1d_params,
 * src/realtime/camera_models.cpp:228-248): theta for 1024 evenly spaced
 * distorted radii up to the image corner plus 10 pixels, by Newton inversion */
rt_status rt_camera_build_lut(rt_camera* camera);

/* The analytic light compiler's derived fields, for lights built by hand
 * (src/scene/analytic_light_compiler.cpp:178-240 and its finalize step,
 * :15-39), in FP64: world_area (sphere 4 pi (r s)^2 and cylinder
 * 2 pi (r s)(l s) with s the basis' similarity scale, disk pi r^2 and rect
 * w h times |basis_x x basis_y|), delta for a distant light whose
 * cos_theta_max is 1, and selection weights (world area, cone solid angle
 * (1 for a delta light) or 4 pi, times Rec.709 luminance) normalized into
 * selection_pdf and a cdf closed at 1 from the last light with positive
 * probability.  A sphere or cylinder whose basis is not a similarity
 * transform is refused, as the reference refuses it. */
rt_status rt_analytic_lights_compile(rt_analytic_light* lights, int32_t count);

/* PackedCameraRig, src/realtime/camera_rig.h:31-34 */
typedef struct rt_camera_rig {
    int32_t active_count;
    int32_t pad;
    rt_camera cameras[4];
} rt_camera_rig;

/* RenderProfile, src/realtime/render_profile.h:10-25, plus two extensions */
typedef struct rt_profile {
    int32_t samples_per_pixel, max_bounces, enable_denoise, rr_start_bounce;
    double accumulation_reset_rotation_deg, accumulation_reset_translation;
    int32_t enable_restir_di, restir_initial_candidates, restir_temporal_reuse,
        restir_max_history_age, restir_max_temporal_candidates, restir_spatial_neighbors,
        restir_max_spatial_candidates, restir_bias_correction, restir_min_analytic_lights;
    int32_t pixel_jitter;          /* extension: jitter primary rays inside the pixel */
    int32_t restir_spatial_radius; /* extension: spatial reuse radius in pixels */
    int32_t pad0;
} rt_profile;

/* ------------------------------------------------------------------ BVH
 * Binary Morton LBVH (Karras 2012) over all primitives.  Node = 64 bytes:
 * the boxes of both children followed by the two child references, the
 * parent index and padding.  A reference with bit 31 set is a leaf whose low
 * 31 bits index the Morton-sorted primitive order; 0xFFFFFFFF is empty. */
typedef struct rt_bvh_node {
    float lo0[3]; float hi0[3];
    float lo1[3]; float hi1[3];
    uint32_t child[2];
    uint32_t parent;
    uint32_t pad;
} rt_bvh_node;

typedef struct rt_bvh_stats {
    int32_t prim_count;
    int32_t node_count;      /* binary LBVH internal nodes */
    int32_t max_depth;       /* binary LBVH depth */
    int32_t wide_node_count; /* nodes of the 4-wide tree every ray traverses */
    double build_ms;         /* device time of the build kernels (LBVH + collapse) */
    int32_t wide_levels;     /* levels of the 4-wide tree */
    int32_t pad;
} rt_bvh_stats;

/* GpuSceneAcceleration::update's classification
 * (src/realtime/gpu/gpu_scene_acceleration.cpp:247-291): reuse (nothing
 * changed), update (materials, lights, background, textures or per-vertex
 * attributes changed; the BVH is kept), refit (primitive positions changed,
 * counts did not: node bounds recomputed bottom-up over the same Morton order
 * and topology, then the 4-wide tree re-collapsed), rebuild (primitive counts
 * changed, or no BVH yet). */
typedef enum rt_acceleration_update_kind {
    RT_ACCEL_REUSE = 0,
    RT_ACCEL_UPDATE = 1,
    RT_ACCEL_REFIT = 2,
    RT_ACCEL_REBUILD = 3
} rt_acceleration_update_kind;

/* AccelerationUpdateStats (gpu_scene_acceleration.h:21-30) */
typedef struct rt_acceleration_update {
    int32_t kind;            /* rt_acceleration_update_kind */
    int32_t node_count;      /* binary LBVH nodes */
    int32_t primitive_reference_count;
    int32_t wide_node_count;
    double elapsed_ms;       /* host time of the whole update (uploads, BVH work, waits) */
    double bvh_ms;           /* device time of the BVH kernels (0 for reuse / update) */
    uint64_t generation;     /* incremented by every update, refit and rebuild */
} rt_acceleration_update;

/* 4-wide quantized BVH node (64 B), collapsed from the binary LBVH: each node
 * absorbs binary descendants, largest surface area first, until it has 4
 * children; binary subtrees of at most 2 primitives become leaf ranges.
 * Child k's box is origin + [qlo, qhi] * 2^(e - 127) per axis (e = byte a of
 * meta, qlo/qhi = byte k of the axis word), rounded outward.  Child
 * reference: 0x80000000 | (count - 1) << 28 | first sorted leaf slot for a
 * leaf range, a node index otherwise, 0xFFFFFFFF when empty. */
typedef struct rt_bvh_wide_node {
    float origin[3];
    uint32_t meta;   /* ex | ey << 8 | ez << 16 | child count << 24 */
    uint32_t child[4];
    uint32_t qlo[3];
    uint32_t qhi[3];
    uint32_t pad[2];
} rt_bvh_wide_node;

typedef struct rt_ray { float o[3]; float tmin; float d[3]; float tmax; } rt_ray;
typedef struct rt_hit { float t; int32_t prim; float u; float v; } rt_hit;

/* ---------------------------------------------------------------- frames */
/* RadianceTiming, src/realtime/gpu/optix_renderer.h:26-30, plus GPU counters */
typedef struct rt_frame_timing {
    double render_ms;      /* device time of the whole frame (all cameras) */
    double denoise_ms;
    double download_ms;
    double extend_ms;      /* traversal launches, summed: each bounce's closest-hit rays together
                              with the previous bounce's shadow (any-hit) rays */
    double shade_ms;
    double connect_ms;     /* the last bounce's shadow rays (their own launch) */
    double other_ms;       /* raygen + accumulate + resolve */
    uint64_t paths;        /* camera paths started */
    uint64_t extend_rays;  /* closest-hit rays traced */
    uint64_t shadow_rays;  /* any-hit rays traced */
    uint64_t extend_launches;
    uint64_t node_visits;  /* closest-hit BVH child boxes tested, one canonical node each
                              (count_traversal only) */
    uint64_t prim_tests;   /* closest-hit primitive tests (count_traversal only) */
    uint64_t sphere_tests; /* of which sphere tests */
    uint64_t shadow_node_visits; /* any-hit BVH child boxes tested (count_traversal only) */
    uint64_t shadow_prim_tests;  /* any-hit primitive tests (count_traversal only) */
    uint64_t shadow_sphere_tests;/* of which sphere tests */
} rt_frame_timing;

/* DeviceRadianceFrameView, src/realtime/gpu/device_radiance_frame_view.h */
typedef struct rt_device_frame_view {
    int32_t width, height;
    const float* beauty; /* device float4 per pixel (denoised when the profile asks for it) */
    const float* normal; /* 0.5 n + 0.5 encoding, temporally resolved */
    const float* albedo;
    const float* depth;  /* device float per pixel */
    const float* resolved_beauty; /* float4: beauty after reprojection, before denoising */
    const float* flow;   /* float2: pixel motion to the previous frame */
    const float* trust;  /* float: 1 where the history texel was accepted */
} rt_device_frame_view;

/* RestirDiagnostics (src/realtime/gpu/optix_renderer.h:40-48, computed as
 * optix_renderer.cpp:183-215 does): over one camera's reservoirs of the last
 * frame, the valid ones, those with temporal / spatial candidates or bias
 * correction sources, and the largest candidate count and age */
typedef struct rt_restir_stats {
    int32_t pixel_count;
    int32_t active_reservoir_count;
    int32_t temporal_reuse_count;
    int32_t spatial_reuse_count;
    int32_t bias_corrected_count;
    int32_t max_candidate_count;
    int32_t max_age;
    int32_t pad;
} rt_restir_stats;

typedef struct rt_device rt_device;
typedef struct rt_scene rt_scene;
typedef struct rt_renderer rt_renderer;

/* ------------------------------------------------------------------- API */
int rt_abi_version(void);
/* thread-local message of the last failing call on this thread */
const char* rt_last_error(void);

/* replaces the device set-up of OptixRenderer::OptixRenderer
 * (src/realtime/gpu/optix_renderer.cpp:118-138) */
rt_status rt_device_create(int ordinal, rt_device** out);
rt_status rt_device_destroy(rt_device* dev);
rt_status rt_device_synchronize(rt_device* dev);
/* name, CU count, HBM bytes; any pointer may be NULL */
/* the device's stream (a hipStream_t): every call on the device is
 * enqueued on it; DeviceRadianceFrameView::stream
 * (device_radiance_frame_view.h:7-12) */
rt_status rt_device_stream(rt_device* dev, void** stream);
rt_status rt_device_info(rt_device* dev, char* name, size_t name_len, int* cu_count,
    uint64_t* total_bytes, uint64_t* free_bytes);

/* replaces prepare_gpu_scene + DeviceSceneBuffers::upload
 * (src/realtime/gpu/packed_scene_preparation.cpp:306-348,
 *  src/realtime/gpu/device_scene_buffers.cpp:58-95) */
rt_status rt_scene_upload(rt_device* dev, const rt_scene_desc* desc, rt_scene** out);
rt_status rt_scene_destroy(rt_scene* scene);

/* replaces GpuSceneAcceleration::rebuild (the CPU median split of
 * src/realtime/gpu/gpu_scene_acceleration.cpp:145-179,314-346) with a GPU LBVH */
rt_status rt_bvh_build(rt_scene* scene, rt_bvh_stats* stats);
/* GpuSceneAcceleration::update (gpu_scene_acceleration.cpp:247-291) and
 * SharedGpuSceneState::prepare (optix_renderer.h:57-66): replaces the scene's
 * contents with `desc`, classified as above by FNV-1a 64 signatures of the
 * primitive arrays and of the whole description, and brings the BVH up to
 * date (rebuild / refit / kept).  Renderers that share the scene see the new
 * contents from their next frame. */
rt_status rt_scene_update(rt_scene* scene, const rt_scene_desc* desc, rt_acceleration_update* out);
/* the last update's statistics (the initial upload + rt_bvh_build counts as a rebuild) */
rt_status rt_scene_last_update(rt_scene* scene, rt_acceleration_update* out);
/* copies node_count nodes, prim_count leaf primitive ids and Morton codes;
 * any pointer may be NULL */
rt_status rt_bvh_download(rt_scene* scene, rt_bvh_node* nodes, uint32_t* sorted_prims,
    uint32_t* morton);
/* copies stats.wide_node_count 4-wide nodes (rt_bvh_wide_node) */
rt_status rt_bvh_download_wide(rt_scene* scene, void* nodes4);

/* test hooks replacing intersect_scene / is_occluded
 * (src/realtime/gpu/programs.cu:1654-1762); host arrays of n rays */
rt_status rt_trace_closest(rt_scene* scene, const rt_ray* rays, int64_t n, rt_hit* hits);
rt_status rt_trace_any(rt_scene* scene, const rt_ray* rays, int64_t n, int32_t* occluded);
/* device-resident variant for throughput measurement: d_rays/d_hits are
 * device pointers; kernel_ms receives the traversal kernel time */
rt_status rt_trace_closest_device(rt_scene* scene, const rt_ray* d_rays, int64_t n,
    rt_hit* d_hits, double* kernel_ms);
/* Allocation-free variant of rt_trace_closest_device, with no counterpart in
 * the reference: for callers that cannot allocate (code under HIP graph
 * capture, serving loops).  It calls no hipMalloc / hipFree and no
 * allocating library routine: its fetch cursor and traversal stack spill
 * live in a file-scope device array of 16 MiB + 256 B per device, resolved
 * at rt_device_create; the persistent grid is cut to what that array holds
 * (a scene whose per-block stack spill exceeds it is refused with
 * RT_ERROR_INVALID_ARGUMENT).  Enqueued on the device's stream without
 * waiting (rt_device_synchronize).  Every call on a device shares the one
 * array: calls that overlap on one device race. */
rt_status rt_trace_closest_device_noalloc(rt_scene* scene, const rt_ray* d_rays, int64_t n,
    rt_hit* d_hits);

/* replaces OptixRenderer + RendererPool (src/realtime/gpu/optix_renderer.h:69-128,
 * src/realtime/gpu/renderer_pool.h:30-53): one renderer traces every
 * camera of a rig in one wavefront pipeline */
rt_status rt_renderer_create(rt_device* dev, rt_scene* scene, rt_renderer** out);
rt_status rt_renderer_destroy(rt_renderer* r);
/* replaces render_prepared_device / launch_radiance_pipeline
 * (src/realtime/gpu/optix_renderer.cpp:230-277,434-457) */
rt_status rt_render_frame(rt_renderer* r, const rt_camera_rig* rig, const rt_profile* profile,
    uint32_t sample_stream, rt_frame_timing* timing);
/* rt_render_frame with each camera slot's random-number identity given:
 * slot c draws from sample stream sample_streams[c] (four values) under the
 * camera key camera_keys[c] (four values in [0, 4); NULL: the slot index).
 * The reference keeps one renderer, and one launch_sample_stream_ counter,
 * per camera (optix_renderer.h:127, renderer_pool.cpp:160-185): per-slot
 * streams carry those counters into one launch, and a camera rendered alone
 * in slot 0 with key k draws the samples of camera k of a whole-rig frame.
 * rt_render_frame(..., s, ...) equals streams {s, s, s, s}, keys {0, 1, 2, 3}. */
rt_status rt_render_frame_streams(rt_renderer* r, const rt_camera_rig* rig,
    const rt_profile* profile, const uint32_t* sample_streams, const uint32_t* camera_keys,
    rt_frame_timing* timing);
/* No counterpart in the reference (whose render_prepared_device waits for
 * its launch): with async on, an rt_render_frame* call whose timing is NULL
 * returns once the frame is enqueued on the device's stream, so the host
 * prepares frame k + 1 while frame k runs.  The renderer's host staging is
 * double-buffered: a call waits only for the frame two calls back.  Its
 * outputs are complete in stream order (every download here waits for the
 * stream; a device view is read on the device's stream or after
 * rt_renderer_synchronize).  A call with timing always waits for its frame.
 * The radiance cache's record / query counts of frames not yet waited on
 * are folded in by rt_renderer_synchronize (or a later frame).  Off by
 * default; turning it off waits for pending frames. */
rt_status rt_renderer_set_async(rt_renderer* r, int32_t enable);
/* waits for every frame of this renderer enqueued on the device's stream */
rt_status rt_renderer_synchronize(rt_renderer* r);
/* render_direction_debug_frame / OptixRenderer::render_direction_debug
 * (direction_debug_renderer.cpp:51-66, programs.cu:1807-1826): the world
 * direction of camera camera_index's ray through each pixel centre, encoded
 * as RGBA8 (255 * clamp01(0.5 (d + 1)) truncated per axis, alpha 255) into
 * `rgba` (width * height * 4 bytes of host memory); waits for the device */
rt_status rt_render_direction_debug(rt_device* dev, const rt_camera_rig* rig, int camera_index,
    uint8_t* rgba);
/* LaunchParameterDiagnostics (optix_renderer.h:50-53, optix_renderer.cpp:
 * 137,217-219,257): the per-frame parameter block is allocated once per
 * renderer (allocation_count 1) and uploaded once per frame launched */
rt_status rt_renderer_launch_diagnostics(rt_renderer* r, uint64_t* allocation_count,
    uint64_t* upload_count);
/* replaces download_radiance_frame_profiled (optix_renderer.cpp:283-326):
 * W*H*4 floats for beauty/normal/albedo, W*H floats for depth; NULL skips */
rt_status rt_frame_download(rt_renderer* r, int camera, float* beauty_rgba, float* normal_rgba,
    float* albedo_rgba, float* depth);
/* the denoiser guides of launch_resolve_kernel (programs.cu:1965-2045): flow
 * (2 floats per pixel), flow trust (1 float) and the resolved beauty before
 * denoising (4 floats); NULL skips */
rt_status rt_frame_download_guides(rt_renderer* r, int camera, float* flow_xy, float* trust,
    float* resolved_beauty_rgba);
/* device view valid until the next rt_render_frame on this renderer */
rt_status rt_frame_device_view(rt_renderer* r, int camera, rt_device_frame_view* view);
/* OptixRenderer::reset_accumulation / reset_sequence.  The sample stream is
 * an argument of every rt_render_frame* call here (the C++ Renderer keeps
 * the reference's per-launch counter); rt_reset_sequence records the seed
 * and resets accumulation. */
rt_status rt_reset_accumulation(rt_renderer* r);
rt_status rt_reset_sequence(rt_renderer* r, uint32_t seed);
rt_status rt_restir_diagnostics(rt_renderer* r, int camera, rt_restir_stats* out);
/* the last frame's RestirReservoir array of one camera (W*H x 84 B,
 * src/common/restir_di.h:36-48 layout); replaces the reservoir download of
 * OptixRenderer::restir_diagnostics (optix_renderer.cpp:183-215) */
rt_status rt_restir_download(rt_renderer* r, int camera, void* reservoirs);

/* ------------------------------------------------- viewer presentation
 * replaces CudaGlTexturePresenter (src/realtime/viewer/cuda_gl_presenter.h:
 * 14-40, cuda_gl_presenter.cu:45-130).  The reference tone maps a device
 * beauty frame into an OpenGL pixel-unpack buffer and uploads it to a
 * texture; with no GL interop here the presenter owns a device RGBA8 buffer
 * of width x height pixels (4 bytes each, alpha 255) that the caller reads
 * or hands to its own display path.  rt_presenter_present reports the
 * CudaGlPresentStatus outcome in *present_status: RT_OK with a present
 * status other than RT_PRESENT_OK is a refused frame, not a failed call. */
typedef enum rt_present_status {
    RT_PRESENT_OK = 0,
    RT_PRESENT_INVALID_DEVICE_VIEW = 1, /* view->beauty is NULL */
    RT_PRESENT_RESOLUTION_MISMATCH = 2  /* view size differs from the presenter's */
} rt_present_status;
typedef struct rt_presenter rt_presenter;
/* the constructor (cuda_gl_presenter.cu:48-67): width, height > 0 */
rt_status rt_presenter_create(rt_device* dev, int32_t width, int32_t height, rt_presenter** out);
rt_status rt_presenter_destroy(rt_presenter* p);
/* CudaGlTexturePresenter::present (cuda_gl_presenter.cu:86-130): tone maps
 * view->beauty (device float4 per pixel, e.g. from rt_frame_device_view)
 * with linear_to_display_u8 into the presenter's buffer on the device's
 * stream and waits for it */
rt_status rt_presenter_present(rt_presenter* p, const rt_device_frame_view* view,
    int32_t* present_status);
/* device address of the presenter's RGBA8 buffer, and a copy to host memory
 * (width * height * 4 bytes) */
rt_status rt_presenter_pixels(rt_presenter* p, uint8_t** d_rgba8);
rt_status rt_presenter_download(rt_presenter* p, uint8_t* rgba8);
/* tone_map_to_rgba8 (cuda_gl_presenter.cu:28-38) between caller buffers in
 * device memory: pixel_count float4 in, pixel_count RGBA8 words out;
 * enqueued on the device's stream without waiting (rt_device_synchronize) */
rt_status rt_tone_map_rgba8(rt_device* dev, const float* d_beauty_rgba, uint8_t* d_rgba8,
    int64_t pixel_count);

/* ------------------------------------------------------ learned denoiser
 * A 4-level convolutional encoder-decoder (widths 32 / 48 / 64 / 96, 15 3x3
 * convolutions with ReLU, mean-pool down, nearest up, skip connections, a
 * 1x1 FP32 head predicting log radiance as a residual of log1p(beauty);
 * 565,443 parameters) run by BF16 MFMA kernels with FP32 accumulation and
 * trained by the same kernels (backward + Adam).  Input features: 1-spp
 * beauty, albedo, normal and depth.  Replaces the closed OptiX denoiser of
 * OptixDenoiserWrapper::run (src/realtime/gpu/denoiser.cpp:111-185).
 * Frames: float4 beauty, 0.5 n + 0.5 encoded float4 normal, float4 albedo,
 * float depth per pixel; the images of a batch lie back to back. */
typedef struct rt_nn rt_nn;
typedef struct rt_nn_images {
    int32_t count;        /* 1..16 */
    int32_t offset[17];   /* first pixel of image i; offset[count] = total */
    int32_t width[16];
} rt_nn_images;
int32_t rt_nn_param_count(void);
rt_status rt_nn_create(int ordinal, rt_nn** out);   /* parameters start at 0 */
rt_status rt_nn_destroy(rt_nn* nn);
/* FP32 parameters: per convolution W[cout][tap][cin] then b[cout]; head W[3][32], b[3].
 * Setting them resets the optimiser state. */
rt_status rt_nn_set_params(rt_nn* nn, const float* params, int32_t n);
rt_status rt_nn_get_params(rt_nn* nn, float* params, int32_t n);
/* device pointers */
rt_status rt_nn_infer(rt_nn* nn, const rt_nn_images* images, const float* beauty,
    const float* normal, const float* albedo, const float* depth, float* out);
/* one training step on a batch: loss = mean over pixels and RGB of
 * w (out_log - log1p(t))^2, t = max(target, 0), w = (1 + t)^2 / (t^2 + 1e-2); the
 * gradient (before the update, host array of rt_nn_param_count floats) and
 * the loss are returned when non-NULL; lr = 0 computes them without updating */
rt_status rt_nn_train_step(rt_nn* nn, const rt_nn_images* images, const float* beauty,
    const float* normal, const float* albedo, const float* depth, const float* target, float lr,
    float* loss, float* grad);
/* the same on host arrays (uploads, runs, downloads) */
rt_status rt_nn_infer_host(rt_nn* nn, const rt_nn_images* images, const float* beauty,
    const float* normal, const float* albedo, const float* depth, float* out);
rt_status rt_nn_train_step_host(rt_nn* nn, const rt_nn_images* images, const float* beauty,
    const float* normal, const float* albedo, const float* depth, const float* target, float lr,
    float* loss, float* grad);
/* Tile shapes of the network's convolution kernels (no counterpart in the
 * reference): rows of output pixels per block at each of the 4 levels, 4 or
 * 8 (8 re-reads fewer halo rows, 4 needs less LDS per block); 0 picks 8 when
 * the level's widest image is at least 160 pixels.  rt_nn_autotune times
 * `reps` inferences of the image batch per candidate (coordinate descent
 * over the levels, median after a warm-up) and applies the fastest. */
typedef struct rt_nn_tiling {
    int32_t conv_rows[4];
} rt_nn_tiling;
typedef struct rt_nn_autotune_result {
    rt_nn_tiling best;
    double best_ms;        /* median inference time of the chosen tiling */
    double baseline_ms;    /* median inference time before the search */
    int32_t trials;
    int32_t pad;
} rt_nn_autotune_result;
rt_status rt_nn_set_tiling(rt_nn* nn, const rt_nn_tiling* t);
rt_status rt_nn_get_tiling(rt_nn* nn, rt_nn_tiling* t);
rt_status rt_nn_autotune(rt_nn* nn, const rt_nn_images* images, int32_t reps,
    rt_nn_autotune_result* out);
/* denoise with `nn` instead of the a-trous filter when the profile asks for
 * denoising (NULL detaches); nn must live on the renderer's device.  The
 * renderer holds its own reference, so rt_nn_destroy on the caller's handle
 * is allowed while the network is attached. */
rt_status rt_renderer_set_denoiser(rt_renderer* r, rt_nn* nn);

/* The learned denoiser's fallback (no counterpart in the reference, whose
 * OptiX denoiser has none): after the network, every camera's output is
 * checked on the device over its surface pixels, and a camera whose output
 * has a non-finite value, moves the mean of log(1 + luminance) by more
 * than max_mean_shift (relative), or keeps more than max_residual_noise of the
 * input's high-frequency energy (sum of |L(p) - L(right neighbour)|) is
 * denoised by the a-trous reference filter instead, in the same frame,
 * without a host round trip (graph replays included).  Defaults 0.8 / 0.9. */
rt_status rt_renderer_set_denoiser_budget(rt_renderer* r, float max_mean_shift,
    float max_residual_noise);
typedef struct rt_denoiser_guard {
    int32_t fallback[4];      /* 1: the camera's last frame came from the filter */
    float nonfinite[4];       /* non-finite network outputs (surface pixels) */
    float mean_shift[4];      /* |mean log1p L(out) - mean log1p L(in)| / mean log1p L(in) */
    float residual_noise[4];  /* HF(out) / HF(in) */
    uint32_t frames_checked;  /* frames the guard ran on */
    uint32_t frames_fallback; /* camera-frames that fell back */
    float max_mean_shift, max_residual_noise;
} rt_denoiser_guard;
/* the guard's state after the last frame (synchronises the device's stream) */
rt_status rt_renderer_denoiser_guard(rt_renderer* r, rt_denoiser_guard* out);

/* tuning knobs chosen by the autotuner; 0 keeps the current value */
typedef struct rt_tuning {
    int32_t wavefront_paths;  /* paths in flight per wavefront (queue capacity) */
    int32_t trace_block;      /* threads per block of the traversal kernels */
    int32_t trace_blocks_per_cu; /* persistent blocks per CU */
    int32_t use_graph;        /* capture the per-frame stage sequence in a hipGraph */
    int32_t count_traversal;  /* count node visits / primitive tests (slower; for measurement) */
    int32_t trace_refill;     /* idle lanes (1..64) before a wave takes new rays; 64 = whole batches */
} rt_tuning;
rt_status rt_renderer_set_tuning(rt_renderer* r, const rt_tuning* t);
rt_status rt_renderer_get_tuning(rt_renderer* r, rt_tuning* t);

/* Autotuner: coordinate descent over traversal block shape, persistent blocks
 * per CU, refill threshold, paths in flight and graph capture, each candidate
 * timed on `frames_per_trial` frames of this rig and profile (median device
 * time after one warm-up frame); with a learned denoiser attached and
 * denoising on, then the network's convolution tile shapes for the rig's
 * camera images (rt_nn_autotune, `denoiser`).  Applies the best tuning and
 * resets the accumulation history. */
typedef struct rt_autotune_result {
    rt_tuning best;
    double best_ms;        /* median frame time of the chosen tuning */
    double baseline_ms;    /* median frame time of the tuning before the search */
    int32_t trials;
    int32_t denoiser_tuned; /* 1: `denoiser` holds the network's search */
    rt_nn_autotune_result denoiser;
} rt_autotune_result;
rt_status rt_renderer_autotune(rt_renderer* r, const rt_camera_rig* rig, const rt_profile* p,
    int32_t frames_per_trial, rt_autotune_result* out);

/* ------------------------------------------------- radiance cache
 * Online radiance cache (no counterpart in the reference: the north-star
 * learned part that ends paths early).  A multiresolution hash-grid encoding
 * of the hit position feeds a 64-wide MLP (four hidden ReLU layers, RGB
 * output) together with the outgoing direction, shading normal, roughness and
 * albedo; inference and training run as BF16 MFMA kernels with FP32
 * accumulation.  Attached to a renderer, the cache is trained every frame on
 * path suffixes that a random subset of paths (training paths) traces with
 * the plain integrator from vertex `query_bounce` on, and once it is usable
 * every other path ends at that vertex with throughput * cache prediction.
 * The plain tracer is the fallback: the renderer does not end paths with the
 * cache before `warmup_frames` frames of training or while the running
 * relative loss on fresh training samples exceeds `max_rel_loss`. */
typedef struct rt_cache_config {
    int32_t levels;           /* hash-grid levels, 1..16 */
    int32_t log2_table;       /* entries per level = 2^log2_table, 8..22 */
    int32_t base_resolution;  /* grid resolution of the first level (>= 2) */
    int32_t max_resolution;   /* grid resolution of the last level */
    int32_t train_samples;    /* training records per frame (capacity), 256..1048576 */
    int32_t train_steps;      /* Adam steps per frame (the records split into this many batches) */
    int32_t query_bounce;     /* path vertex where paths end with the cache (>= 1) */
    int32_t warmup_frames;    /* frames of training before the cache ends paths */
    float learning_rate;      /* Adam step size for the MLP and the grid */
    float max_rel_loss;       /* fallback threshold on the running relative loss (its
                                 fresh single-sample targets keep it near 1-4 on a
                                 working cache; the default 8 catches divergence) */
    float train_fraction;     /* share of pixel samples traced fully as training paths;
                                 0: adaptive, from 1.5 x train_samples / paths of the
                                 frame toward 1.1 x train_samples records a frame */
    uint32_t seed;            /* parameter initialisation */
} rt_cache_config;

typedef struct rt_cache_stats {
    float loss_ema;           /* running relative loss on fresh training samples */
    float last_loss;          /* relative loss of the last frame's first batch, before its step */
    int32_t frames;           /* frames trained */
    int32_t usable;           /* the renderer ends paths with the cache */
    int32_t steps;            /* Adam steps taken */
    int32_t records;          /* training records of the last frame */
    int32_t queries;          /* paths ended with the cache in the last frame */
    float train_fraction;     /* share of pixel samples traced as training paths next */
} rt_cache_stats;

typedef struct rt_cache rt_cache;
/* 16 levels, 2^19 entries, resolutions 16..2048, 65536 samples in 4 steps,
 * vertex 2, 4 warm-up frames, learning rate 1e-2, automatic training share */
void rt_cache_default_config(rt_cache_config* c);
int32_t rt_cache_mlp_param_count(void);
/* MLP parameters then the grid (levels x 2^log2_table x 2) */
int64_t rt_cache_param_count(const rt_cache_config* c);
rt_status rt_cache_create(int ordinal, const rt_cache_config* c, rt_cache** out);
rt_status rt_cache_destroy(rt_cache* cache);
/* FP32 parameters: W0..W3 [64 out][64 in], W4 [3][64], grid [L][T][2].
 * Setting them resets the optimiser and the training state. */
rt_status rt_cache_set_params(rt_cache* cache, const float* params, int64_t n);
rt_status rt_cache_get_params(rt_cache* cache, float* params, int64_t n);
/* the box the hash grid spans (the renderer sets its scene's bounds) */
rt_status rt_cache_set_bounds(rt_cache* cache, const float lo[3], const float hi[3]);
/* the encoding's exact constants: per-level resolution and dense flag, box
 * origin and reciprocal extent */
rt_status rt_cache_grid_info(rt_cache* cache, float res[16], int32_t dense[16], float lo[3],
    float inv_extent[3]);
/* inputs: n x 16 floats (pos.xyz, 0, outgoing dir.xyz, roughness,
 * normal.xyz, 0, albedo.rgb, 0); out: n x 4 floats (RGB, 0) */
rt_status rt_cache_infer_host(rt_cache* cache, const float* inputs, int32_t n, float* out);
/* one Adam step on n samples (targets n x 4 floats, RGB used); loss = the
 * relative loss before the step; grad (optional, rt_cache_param_count
 * floats) = its gradient */
rt_status rt_cache_train_step_host(rt_cache* cache, const float* inputs, const float* targets,
    int32_t n, float* loss, float* grad);
rt_status rt_cache_get_stats(rt_cache* cache, rt_cache_stats* out);
/* trains and uses `cache` in this renderer's frames (NULL detaches); the
 * renderer holds its own reference */
rt_status rt_renderer_set_radiance_cache(rt_renderer* r, rt_cache* cache);

/* ------------------------------------------------- synthetic scenes
 * Seeded generators of the benchmark scenes (this repository's equivalent
 * of build_scene(id) + make_rig, src/scene/shared_scene_builders.h:56-57):
 *   0 random_spheres, 1 cornell_mesh, 2 instanced_mesh_restir,
 *   3 radiance_cache_realtime, 4 cornell_random (seeded denoiser training scenes),
 *   5 smoke, 6 final_room (the reference's builtin realtime scenes of those
 *   names, src/scene/shared_scene_builders.cpp:475-550, with its four-camera
 *   surround rig, src/realtime/viewer/four_camera_rig.cpp:55-95) */
typedef struct rt_generated_scene {
    rt_scene_desc scene;
    int32_t camera_count;
    int32_t pad;
    rt_camera cameras[4];
    rt_profile profile;
} rt_generated_scene;
rt_status rt_scene_generate(int config, uint64_t seed, rt_generated_scene** out);
rt_status rt_scene_generated_free(rt_generated_scene* g);
/* camera rig of frame `frame` of a `frames`-frame orbit (configs 2 and 3; the other
 * configs have static rigs and return their frame-0 rig) */
rt_status rt_scene_generate_rig(int config, uint64_t seed, int frame, int frames,
    rt_camera_rig* out);

#ifdef __cplusplus
}
#endif
#endif
