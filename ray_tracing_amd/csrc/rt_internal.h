// rt_internal.h -- structures shared by the host orchestration and the gfx950
// kernels of the path tracer.
//
// HBM layout (one scene):
//   nodes   : BVH node array, 4 x float4 = 64 B per node (child boxes,
//             child references, parent) -- same bytes as rt_bvh_node
//   leaves  : 3 x float4 = 48 B per primitive in Morton order
//             triangle (p0 | id), (p1-p0), (p2-p0); quad (origin | id),
//             edge_u, edge_v; sphere (center | id), (radius, 0, 0, 0)
//             where id = global primitive id | type << 30
//   lights  : power CDF over emissive primitives + environment, and a
//             primitive -> light index table (O(1) MIS lookups)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rt/rt_hip.h"

namespace rtk {

enum : uint32_t { kPrimSphere = 0, kPrimQuad = 1, kPrimTriangle = 2 };
constexpr uint32_t kIdMask = 0x3FFFFFFFu;

struct LightRec {
    int32_t type;  // 0 sphere, 1 quad, 2 triangle, 3 environment
    int32_t prim;  // index within its type, -1 for the environment
    float selection_pdf;
    float cdf;
};

struct SceneView {
    const float4* nodes;   // binary LBVH (build intermediate, test downloads)
    const float4* leaves;
    int32_t node_count;
    int32_t prim_count;
    const float4* nodes4;  // 4-wide quantized BVH traversed by every ray
    int32_t node4_count;
    int32_t stack4_entries; // deepest deferred-node stack a 4-wide traversal can need
    const rt_sphere* spheres;
    const rt_quad* quads;
    const rt_triangle* triangles;
    int32_t sphere_count, quad_count, triangle_count, material_count;
    const rt_material* materials;
    const rt_openpbr_material* openpbr;
    int32_t openpbr_count;
    int32_t pbr_full;   // a material uses OpenPBR lobes beyond the opaque core
    int32_t light_count;
    const LightRec* lights;
    const int32_t* light_of_prim;
    int32_t env_light;
    int32_t background_mode;
    float bg[3];
    float bg_top[3];
    const rt_analytic_light* analytic;
    int32_t analytic_count;
    int32_t stack_depth;  // LDS traversal stack entries per lane (>= BVH depth)
    const float* a_sel;   // analytic light selection pdf
    const float* a_cdf;   // analytic light CDF
    const int32_t* a_infinite;  // dome / distant light indices, ascending
    int32_t a_infinite_count;
    // binary BVH over the finite, non-point analytic lights in depth-first
    // order: node k is a_nodes[2k] = (lo, skip: the first node after its
    // subtree) and a_nodes[2k + 1] = (hi, light count << 28 | first slot;
    // count 0 for an inner node, whose left child is node k + 1); leaves
    // list lights through a_order
    const float4* a_nodes;
    const int32_t* a_order;
    int32_t a_node_count;
    const rt_medium* media;   // homogeneous media (free-flight events of closest-hit rays)
    int32_t medium_count;
    const rt_texture* textures;
    const float* texels;               // 3 floats per texel
    const int32_t* mat_albedo_tex;     // nullptr: every albedo is a constant
    const float* tri_uv;               // 6 floats per triangle, or nullptr
    const float* tri_n;                // 9 floats per triangle, or nullptr
    int32_t texture_count, texel_count;
    const int32_t* mat_emission_tex;   // nullptr: every emission is a constant
    const rt_openpbr_textures* pbr_tex;  // nullptr: no OpenPBR input is bound to a texture
};

// 4-wide quantized BVH node, 64 B = 4 x float4, collapsed from the binary
// LBVH (each node absorbs binary descendants, largest surface area first,
// until it has 4 children):
//   v0 = origin.xyz (min corner of the children), meta = ex | ey << 8 |
//        ez << 16 | count << 24 (per-axis quantum 2^(e - 127))
//   v1 = child references: a leaf range kLeafBit | (count - 1) << 28 |
//        first sorted leaf slot (1..kLeafMax primitives), a 4-wide node
//        index, or kEmpty
//   v2 = qlo.x, qlo.y, qlo.z, qhi.x   (byte k of each word = child k)
//   v3 = qhi.y, qhi.z, 0, 0
// A child's box is origin + [qlo, qhi] * quantum per axis, rounded outward,
// so it always contains the exact box.
constexpr int kWide = 4;
// binary subtrees with at most this many primitives become one leaf range
#ifndef RT_WIDE_LEAF_MAX
#define RT_WIDE_LEAF_MAX 2
#endif
constexpr uint32_t kLeafMax = RT_WIDE_LEAF_MAX;
__host__ __device__ constexpr uint32_t wide_leaf(uint32_t first, uint32_t count) {
    return 0x80000000u | ((count - 1u) << 28) | first;
}
// LDS part of the per-lane traversal stack; deeper entries spill to HBM
#ifndef RT_STACK_LDS
#define RT_STACK_LDS 16
#endif
constexpr int kStackLds = RT_STACK_LDS;

// RestirReservoir (src/common/restir_di.h:36-48): 84 B per pixel
struct Reservoir {
    int32_t light_index;
    float sample_u0, sample_u1;
    float position[3];
    float normal[3];
    int32_t material_type, primitive_type, primitive_index;
    float weight_sum, selected_target, estimator_weight;
    int32_t candidate_count, temporal_candidate_count, spatial_candidate_count,
        bias_correction_source_count, age, valid;
};

// Values that change from frame to frame, read by the kernels from device
// memory so that a captured graph of the frame's launches stays valid: the
// host uploads them before every frame (eager or graph replay).
struct FrameDyn {
    uint32_t stream[4];         // sample stream of each camera slot (RNG key)
    uint32_t camera_key[4];     // camera key of each slot (RNG key; the slot index by default)
    int32_t restir_history;     // the previous frame ran ReSTIR (its reservoirs are history)
    int32_t history_length[4];  // frames of history per camera (capture_launch_history)
    int32_t prev_valid[4];      // previous camera valid per camera
};

// per-frame ReSTIR DI parameters and per-camera history
struct RestirView {
    int32_t enabled;          // profile + light-count rule of radiance_launch_setup.cpp:101-106
    int32_t initial_candidates;
    int32_t temporal_reuse;
    int32_t max_history_age;
    int32_t max_temporal_candidates;
    int32_t spatial_neighbors;
    int32_t max_spatial_candidates;
    int32_t bias_correction;
    int32_t spatial_radius;
    int32_t spp;
    Reservoir* current;             // all cameras, indexed by global pixel
    const Reservoir* history;       // previous frame, same indexing (valid if dyn->restir_history)
    const rt_camera* prev_cams;     // previous camera per slot (device)
    const FrameDyn* dyn;
};

// per-camera constants for ray generation (double camera math as the
// reference's DeviceActiveCamera)
struct CameraView {
    const rt_camera* cams;   // device copy of the rig's cameras
    int32_t count;
    int32_t pixel_offset[5]; // prefix sum of W*H
    int32_t width[4];
};

}  // namespace rtk
