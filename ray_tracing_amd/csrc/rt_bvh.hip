// rt_bvh.hip -- GPU Morton LBVH build for gfx950.
//
// Replaces the reference's host-side recursive median split
// (src/realtime/gpu/gpu_scene_acceleration.cpp:145-179, 314-346; leaf <= 4)
// with a fully parallel build (Karras 2012, "Maximizing Parallelism in the
// Construction of BVHs, Octrees, and k-d Trees", HPG):
//   k_prim_bounds : primitive boxes (1e-5 padding of thin axes as
//                   gpu_scene_acceleration.cpp:43-71) + centroid bounds by
//                   order-preserving integer atomics (exact)
//   k_morton      : 30-bit codes, 10 bits per axis in the centroid box
//   radix sort    : stable LSD, 4-bit digits, wave64 ballot ranking
//   k_karras      : one thread per internal node (radix tree)
//   k_refit       : one thread per leaf climbing with arrival counters;
//                   writes the Morton-ordered leaf records
//   k_depth       : deepest root-leaf path (sizes the LDS traversal stack)
// Every step is integer work or exact min/max, so the node array is
// bit-identical to a sequential CPU build of the same primitives.
#include "rt_internal.h"
#include "rt_kernels.h"
#include "rt_math.h"

namespace rtk {

namespace {

constexpr int kSortBlock = 256;
constexpr int kSortItems = 4;
constexpr int kSortTile = kSortBlock * kSortItems;
constexpr int kWaves = kSortBlock / 64;

struct Box {
    float lo[3];
    float hi[3];
};

__device__ __forceinline__ void box_empty(Box& b) {
    for (int a = 0; a < 3; ++a) { b.lo[a] = __builtin_inff(); b.hi[a] = -__builtin_inff(); }
}
__device__ __forceinline__ void box_point(Box& b, f3 p) {
    b.lo[0] = sel_min(b.lo[0], p.x); b.hi[0] = sel_max(b.hi[0], p.x);
    b.lo[1] = sel_min(b.lo[1], p.y); b.hi[1] = sel_max(b.hi[1], p.y);
    b.lo[2] = sel_min(b.lo[2], p.z); b.hi[2] = sel_max(b.hi[2], p.z);
}
__device__ __forceinline__ void box_pad(Box& b) {
    for (int a = 0; a < 3; ++a) {
        if (b.hi[a] - b.lo[a] < 1e-5f) { b.lo[a] -= 1e-5f; b.hi[a] += 1e-5f; }
    }
}

__device__ __forceinline__ uint32_t ordered(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float unordered(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

__device__ Box prim_box(const SceneView& s, int id) {
    Box b;
    box_empty(b);
    if (id < s.sphere_count) {
        const rt_sphere sp = s.spheres[id];
        const float r = fabsf(sp.radius);
        for (int a = 0; a < 3; ++a) { b.lo[a] = sp.center[a] - r; b.hi[a] = sp.center[a] + r; }
        return b;
    }
    id -= s.sphere_count;
    if (id < s.quad_count) {
        const rt_quad q = s.quads[id];
        const f3 o = ld3(q.origin), u = ld3(q.edge_u), v = ld3(q.edge_v);
        box_point(b, o);
        box_point(b, o + u);
        box_point(b, o + v);
        box_point(b, (o + u) + v);
        box_pad(b);
        return b;
    }
    id -= s.quad_count;
    const rt_triangle t = s.triangles[id];
    box_point(b, ld3(t.p0));
    box_point(b, ld3(t.p1));
    box_point(b, ld3(t.p2));
    box_pad(b);
    return b;
}

__global__ void k_init_bounds(uint32_t* cb) {
    if (threadIdx.x < 3) cb[threadIdx.x] = 0xFFFFFFFFu;       // min keys
    else if (threadIdx.x < 6) cb[threadIdx.x] = 0u;            // max keys
    if (threadIdx.x == 6) cb[6] = 0u;                          // max depth
}

// primitive boxes and the bounds of their centroids: a bounded grid strides
// over the primitives, each block reduces its waves in LDS, and one lane per
// block merges the result with six atomics (min / max are order-free)
constexpr int kBoundsBlock = 256;
constexpr int kBoundsMaxBlocks = 2048;

__global__ __launch_bounds__(kBoundsBlock) void k_prim_bounds(SceneView s, int n, float4* pbox,
    uint32_t* cb) {
    __shared__ uint32_t red[kBoundsBlock / 64][6];
    uint32_t kmin[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
    uint32_t kmax[3] = {0u, 0u, 0u};
    for (int i = blockIdx.x * kBoundsBlock + (int)threadIdx.x; i < n; i += gridDim.x * kBoundsBlock) {
        const Box b = prim_box(s, i);
        pbox[2 * i] = make_float4(b.lo[0], b.lo[1], b.lo[2], 0.0f);
        pbox[2 * i + 1] = make_float4(b.hi[0], b.hi[1], b.hi[2], 0.0f);
        for (int a = 0; a < 3; ++a) {
            const uint32_t k = ordered((b.lo[a] + b.hi[a]) * 0.5f);
            kmin[a] = min(kmin[a], k);
            kmax[a] = max(kmax[a], k);
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        for (int a = 0; a < 3; ++a) {
            kmin[a] = min(kmin[a], (uint32_t)__shfl_xor((int)kmin[a], off));
            kmax[a] = max(kmax[a], (uint32_t)__shfl_xor((int)kmax[a], off));
        }
    }
    const int wave = (int)threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        for (int a = 0; a < 3; ++a) { red[wave][a] = kmin[a]; red[wave][3 + a] = kmax[a]; }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBoundsBlock / 64; ++w) {
            for (int a = 0; a < 3; ++a) {
                kmin[a] = min(kmin[a], red[w][a]);
                kmax[a] = max(kmax[a], red[w][3 + a]);
            }
        }
        for (int a = 0; a < 3; ++a) {
            atomicMin(&cb[a], kmin[a]);
            atomicMax(&cb[3 + a], kmax[a]);
        }
    }
}

inline int bounds_blocks(int n) {
    const int b = (n + kBoundsBlock - 1) / kBoundsBlock;
    return b < kBoundsMaxBlocks ? b : kBoundsMaxBlocks;
}

__device__ __forceinline__ uint32_t expand10(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__device__ __forceinline__ uint32_t quantize10(float c, float lo, float scale) {
    float q = (c - lo) * scale;
    q = q < 0.0f ? 0.0f : q;
    q = q > 1023.0f ? 1023.0f : q;
    return (uint32_t)q;
}

__global__ void k_morton(int n, const float4* pbox, const uint32_t* cb, uint32_t* keys,
    uint32_t* vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float lo[3], scale[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = unordered(cb[a]);
        const float ext = unordered(cb[3 + a]) - lo[a];
        scale[a] = ext > 0.0f ? 1024.0f / ext : 0.0f;
    }
    const float4 l = pbox[2 * i], h = pbox[2 * i + 1];
    const uint32_t qx = quantize10((l.x + h.x) * 0.5f, lo[0], scale[0]);
    const uint32_t qy = quantize10((l.y + h.y) * 0.5f, lo[1], scale[1]);
    const uint32_t qz = quantize10((l.z + h.z) * 0.5f, lo[2], scale[2]);
    keys[i] = (expand10(qx) << 2) | (expand10(qy) << 1) | expand10(qz);
    vals[i] = (uint32_t)i;
}

// ---- stable LSD radix sort, 4-bit digits
__global__ __launch_bounds__(kSortBlock) void k_radix_count(const uint32_t* keys, int n, int shift,
    uint32_t* counts, int nblocks) {
    __shared__ uint32_t hist[16];
    if (threadIdx.x < 16) hist[threadIdx.x] = 0u;
    __syncthreads();
    const int base = blockIdx.x * kSortTile;
    for (int k = 0; k < kSortItems; ++k) {
        const int idx = base + k * kSortBlock + (int)threadIdx.x;
        if (idx < n) atomicAdd(&hist[(keys[idx] >> shift) & 15u], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 16) counts[threadIdx.x * nblocks + blockIdx.x] = hist[threadIdx.x];
}

// exclusive scan of `total` counts by one block
__global__ __launch_bounds__(1024) void k_scan(uint32_t* counts, int total) {
    __shared__ uint32_t part[1024];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0u;
    __syncthreads();
    for (int base = 0; base < total; base += 1024) {
        const int idx = base + (int)threadIdx.x;
        const uint32_t v = idx < total ? counts[idx] : 0u;
        part[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const uint32_t add = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0u;
            __syncthreads();
            part[threadIdx.x] += add;
            __syncthreads();
        }
        if (idx < total) counts[idx] = carry + part[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry += part[1023];
        __syncthreads();
    }
}

// exclusive scan of a long array in three passes: tile sums, one-block
// scan of the tile sums, then each tile scanned from its offset
constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;

__global__ __launch_bounds__(kScanBlock) void k_tile_sums(const uint32_t* in, int total,
    uint32_t* sums) {
    __shared__ uint32_t part[kScanBlock / 64];
    const int base = blockIdx.x * kScanTile;
    uint32_t acc = 0u;
    for (int k = 0; k < kScanItems; ++k) {
        const int idx = base + k * kScanBlock + (int)threadIdx.x;
        if (idx < total) acc += in[idx];
    }
    for (int off = 32; off > 0; off >>= 1) acc += (uint32_t)__shfl_xor((int)acc, off);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0u;
        for (int w = 0; w < kScanBlock / 64; ++w) t += part[w];
        sums[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(kScanBlock) void k_tile_scan(const uint32_t* in, int total,
    const uint32_t* offsets, uint32_t* out) {
    __shared__ uint32_t part[kScanBlock / 64];
    const int base = blockIdx.x * kScanTile + (int)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t t = 0u;
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = base + k < total ? in[base + k] : 0u;
        t += v[k];
    }
    const int lane = (int)threadIdx.x & 63, wave = (int)threadIdx.x >> 6;
    uint32_t x = t;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)x, off);
        if (lane >= off) x += y;
    }
    if (lane == 63) part[wave] = x;
    __syncthreads();
    uint32_t run = offsets[blockIdx.x] + x - t;
    for (int w = 0; w < wave; ++w) run += part[w];
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < total) out[base + k] = run;
        run += v[k];
    }
}

__global__ __launch_bounds__(kSortBlock) void k_radix_scatter(const uint32_t* keys_in,
    const uint32_t* vals_in, uint32_t* keys_out, uint32_t* vals_out, int n, int shift,
    const uint32_t* offsets, int nblocks) {
    __shared__ uint32_t running[16];
    __shared__ uint32_t wave_count[kWaves][16];
    const int lane = (int)threadIdx.x & 63;
    const int wave = (int)threadIdx.x >> 6;
    if (threadIdx.x < 16) running[threadIdx.x] = offsets[threadIdx.x * nblocks + blockIdx.x];
    __syncthreads();
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int base = blockIdx.x * kSortTile;
    for (int k = 0; k < kSortItems; ++k) {
        const int idx = base + k * kSortBlock + (int)threadIdx.x;
        const bool valid = idx < n;
        const uint32_t key = valid ? keys_in[idx] : 0u;
        const uint32_t val = valid ? vals_in[idx] : 0u;
        const uint32_t dig = (key >> shift) & 15u;
        uint32_t rank = 0u;
        for (uint32_t d = 0; d < 16u; ++d) {
            const uint64_t m = __ballot(valid && dig == d);
            if (valid && dig == d) rank = (uint32_t)__popcll(m & lt);
            if (lane == 0) wave_count[wave][d] = (uint32_t)__popcll(m);
        }
        __syncthreads();
        if (valid) {
            uint32_t pos = running[dig] + rank;
            for (int w = 0; w < wave; ++w) pos += wave_count[w][dig];
            keys_out[pos] = key;
            vals_out[pos] = val;
        }
        __syncthreads();
        if (threadIdx.x < 16) {
            uint32_t tot = 0u;
            for (int w = 0; w < kWaves; ++w) tot += wave_count[w][threadIdx.x];
            running[threadIdx.x] += tot;
        }
        __syncthreads();
    }
}

// ---- Karras radix tree
__device__ __forceinline__ int delta(const uint32_t* keys, int n, int i, int j) {
    if (j < 0 || j > n - 1) return -1;
    const uint32_t a = keys[i], b = keys[j];
    if (a == b) return 32 + __clz((int)((uint32_t)i ^ (uint32_t)j));
    return __clz((int)(a ^ b));
}

__global__ void k_karras(const uint32_t* keys, int n, uint4* nodes_u, uint32_t* leaf_parent) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2) {
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    }
    const int j = i + l * d;
    const int dnode = delta(keys, n, i, j);
    int s = 0;
    int div = 2;
    int t = (l + div - 1) / div;
    while (true) {
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
        if (t == 1) break;
        div *= 2;
        t = (l + div - 1) / div;
    }
    const int gamma = i + s * d + (d < 0 ? d : 0);
    const int first = min(i, j);
    const int last = max(i, j);
    uint32_t c0, c1;
    if (first == gamma) {
        c0 = kLeafBit | (uint32_t)gamma;
        leaf_parent[gamma] = (uint32_t)i;
    } else {
        c0 = (uint32_t)gamma;
        nodes_u[4 * gamma + 3].z = (uint32_t)i;
    }
    if (last == gamma + 1) {
        c1 = kLeafBit | (uint32_t)(gamma + 1);
        leaf_parent[gamma + 1] = (uint32_t)i;
    } else {
        c1 = (uint32_t)(gamma + 1);
        nodes_u[4 * (gamma + 1) + 3].z = (uint32_t)i;
    }
    nodes_u[4 * i + 3].x = c0;
    nodes_u[4 * i + 3].y = c1;
    nodes_u[4 * i + 3].w = (uint32_t)(last - first + 1);  // leaves under this node
    if (i == 0) nodes_u[3].z = kEmpty;
}

// A child box passes from the thread that finished the child to the one
// that finishes the parent, which may run on another XCD with its own L2.
// The box is written with agent-scope atomic stores (through to the
// coherence point) and read back with agent-scope atomic loads; the explicit
// vmcnt(0) wait completes the stores before the arrival counter moves.  An
// agent-scope release / acquire fence would give the same order but writes
// back and invalidates the whole L2 of the XCD, twice per tree level.
__device__ __forceinline__ void publish_child_box(float* node_f, int slot, const float lo[3],
    const float hi[3]) {
    uint32_t* p = reinterpret_cast<uint32_t*>(node_f + slot * 6);
    for (int a = 0; a < 3; ++a) {
        __hip_atomic_store(p + a, __float_as_uint(lo[a]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(p + 3 + a, __float_as_uint(hi[a]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

__device__ __forceinline__ float coherent_load(const float* p) {
    return __uint_as_float(__hip_atomic_load(reinterpret_cast<const uint32_t*>(p),
        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

__device__ void write_leaf(const SceneView& s, uint32_t prim, float4* leaf) {
    uint32_t id = prim;
    if ((int)id < s.sphere_count) {
        const rt_sphere sp = s.spheres[id];
        leaf[0] = make_float4(sp.center[0], sp.center[1], sp.center[2],
            __uint_as_float(prim | (kPrimSphere << 30)));
        leaf[1] = make_float4(sp.radius, 0.0f, 0.0f, 0.0f);
        leaf[2] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    id -= (uint32_t)s.sphere_count;
    if ((int)id < s.quad_count) {
        const rt_quad q = s.quads[id];
        leaf[0] = make_float4(q.origin[0], q.origin[1], q.origin[2],
            __uint_as_float(prim | (kPrimQuad << 30)));
        leaf[1] = make_float4(q.edge_u[0], q.edge_u[1], q.edge_u[2], 0.0f);
        leaf[2] = make_float4(q.edge_v[0], q.edge_v[1], q.edge_v[2], 0.0f);
        return;
    }
    id -= (uint32_t)s.quad_count;
    const rt_triangle t = s.triangles[id];
    const f3 p0 = ld3(t.p0);
    const f3 e1 = ld3(t.p1) - p0;
    const f3 e2 = ld3(t.p2) - p0;
    leaf[0] = make_float4(p0.x, p0.y, p0.z, __uint_as_float(prim | (kPrimTriangle << 30)));
    leaf[1] = f4(e1, 0.0f);
    leaf[2] = f4(e2, 0.0f);
}

__global__ void k_refit(SceneView s, int n, const uint32_t* sorted, const float4* pbox,
    const uint32_t* leaf_parent, float4* nodes, uint32_t* flags, float4* leaves) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t prim = sorted[j];
    write_leaf(s, prim, leaves + 3 * j);
    const float4 l4 = pbox[2 * prim], h4 = pbox[2 * prim + 1];
    float lo[3] = {l4.x, l4.y, l4.z};
    float hi[3] = {h4.x, h4.y, h4.z};
    uint32_t child = kLeafBit | (uint32_t)j;
    uint32_t node = leaf_parent[j];
    uint4* nodes_u = reinterpret_cast<uint4*>(nodes);
    while (true) {
        const uint4 refs = nodes_u[4 * node + 3];
        const int slot = refs.x == child ? 0 : 1;
        publish_child_box(reinterpret_cast<float*>(nodes + 4 * node), slot, lo, hi);
        const uint32_t old = atomicAdd(&flags[node], 1u);
        if (old == 0u) return;
        asm volatile("" ::: "memory");   // the sibling's box is read after the counter
        const float* nf = reinterpret_cast<const float*>(nodes + 4 * node);
        float acc_lo[3], acc_hi[3];
        for (int a = 0; a < 3; ++a) { acc_lo[a] = __builtin_inff(); acc_hi[a] = -__builtin_inff(); }
        for (int c = 0; c < 2; ++c) {
            for (int a = 0; a < 3; ++a) {
                acc_lo[a] = sel_min(acc_lo[a], coherent_load(nf + c * 6 + a));
                acc_hi[a] = sel_max(acc_hi[a], coherent_load(nf + c * 6 + 3 + a));
            }
        }
        for (int a = 0; a < 3; ++a) { lo[a] = acc_lo[a]; hi[a] = acc_hi[a]; }
        const uint32_t parent = refs.z;
        if (parent == kEmpty) return;
        child = node;
        node = parent;
    }
}

__global__ void k_single(SceneView s, const uint32_t* sorted, const float4* pbox, float4* nodes,
    float4* leaves, uint32_t* cb) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    cb[6] = 1u;
    write_leaf(s, sorted[0], leaves);
    const float4 l = pbox[0], h = pbox[1];
    const float inf = __builtin_inff();
    nodes[0] = make_float4(l.x, l.y, l.z, h.x);
    nodes[1] = make_float4(h.y, h.z, inf, inf);
    nodes[2] = make_float4(inf, -inf, -inf, -inf);
    reinterpret_cast<uint4*>(nodes)[3] = make_uint4(kLeafBit, kEmpty, kEmpty, 1u);
}

__global__ void k_depth(int n, const uint32_t* leaf_parent, const float4* nodes, uint32_t* cb) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    uint32_t depth = 1u;
    uint32_t node = leaf_parent[j];
    const uint4* nodes_u = reinterpret_cast<const uint4*>(nodes);
    while (true) {
        const uint32_t parent = nodes_u[4 * node + 3].z;
        if (parent == kEmpty) break;
        ++depth;
        node = parent;
    }
    atomicMax(&cb[6], depth);
}

// ---- collapse to the 4-wide quantized BVH
//
// expand(b): the children of 4-wide node rooted at binary node b.  Start
// from b's two children and, while fewer than 4, replace the internal child
// of largest surface area (first one on ties) by its two children, in place.
// Every step reads boxes stored in binary parents, so the result is a pure
// function of the binary node array.
struct Wide {
    uint32_t ref[kWide];
    float lo[kWide][3];
    float hi[kWide][3];
    int count;
};

__device__ __forceinline__ float box_area(const float* lo, const float* hi) {
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return dx * dy + dy * dz + dz * dx;
}

__device__ void load_binary_children(const float4* nodes, uint32_t b, uint32_t* ref, float (*lo)[3],
    float (*hi)[3]) {
    const float* nf = reinterpret_cast<const float*>(nodes + 4u * b);
    const uint4 refs = reinterpret_cast<const uint4*>(nodes)[4u * b + 3u];
    ref[0] = refs.x;
    ref[1] = refs.y;
    for (int c = 0; c < 2; ++c) {
        for (int a = 0; a < 3; ++a) {
            lo[c][a] = nf[c * 6 + a];
            hi[c][a] = nf[c * 6 + 3 + a];
        }
    }
}

__device__ __forceinline__ uint32_t subtree_leaves(const float4* nodes, uint32_t ref) {
    return (ref & kLeafBit) ? 1u : reinterpret_cast<const uint4*>(nodes)[4u * ref + 3u].w;
}

// leftmost leaf slot under a binary reference
__device__ __forceinline__ uint32_t first_leaf(const float4* nodes, uint32_t ref) {
    while (!(ref & kLeafBit)) ref = reinterpret_cast<const uint4*>(nodes)[4u * ref + 3u].x;
    return ref & ~kLeafBit;
}

__device__ void expand_wide(const float4* nodes, uint32_t b, Wide& w) {
    uint32_t r2[2];
    float l2[2][3], h2[2][3];
    load_binary_children(nodes, b, r2, l2, h2);
    w.count = 0;
    for (int c = 0; c < 2; ++c) {
        if (r2[c] == kEmpty) continue;
        w.ref[w.count] = r2[c];
        for (int a = 0; a < 3; ++a) { w.lo[w.count][a] = l2[c][a]; w.hi[w.count][a] = h2[c][a]; }
        ++w.count;
    }
    while (w.count < kWide) {
        int best = -1;
        float best_area = -1.0f;
        for (int j = 0; j < w.count; ++j) {
            if (subtree_leaves(nodes, w.ref[j]) <= kLeafMax) continue;  // becomes a leaf range
            const float ar = box_area(w.lo[j], w.hi[j]);
            if (ar > best_area) { best_area = ar; best = j; }
        }
        if (best < 0) break;
        load_binary_children(nodes, w.ref[best], r2, l2, h2);
        for (int j = w.count; j > best + 1; --j) {
            w.ref[j] = w.ref[j - 1];
            for (int a = 0; a < 3; ++a) { w.lo[j][a] = w.lo[j - 1][a]; w.hi[j][a] = w.hi[j - 1][a]; }
        }
        for (int c = 0; c < 2; ++c) {
            w.ref[best + c] = r2[c];
            for (int a = 0; a < 3; ++a) { w.lo[best + c][a] = l2[c][a]; w.hi[best + c][a] = h2[c][a]; }
        }
        ++w.count;
    }
}

// one level of the top-down walk: every frontier node marks its internal
// children as 4-wide nodes and appends them to the next frontier
// one level of the top-down collapse; the frontier size comes from device
// memory (grid-stride over it), so levels are enqueued without the host
// reading each size, and a launch over an empty frontier returns at once.
// *levels counts the non-empty frontiers.
__global__ void k_wide_level(const float4* nodes, const uint32_t* front_in, const uint32_t* n_in,
    uint32_t* front_out, uint32_t* n_out, uint32_t* is_wide, uint32_t* levels) {
    const uint32_t n = *n_in;
    if (blockIdx.x == 0 && threadIdx.x == 0 && n > 0u) *levels += 1u;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        Wide w;
        expand_wide(nodes, front_in[i], w);
        for (int k = 0; k < w.count; ++k) {
            if (subtree_leaves(nodes, w.ref[k]) <= kLeafMax) continue;
            is_wide[w.ref[k]] = 1u;
            front_out[atomicAdd(n_out, 1u)] = w.ref[k];
        }
    }
}

__global__ void k_wide_root(uint32_t* front, uint32_t* counts, uint32_t* is_wide) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    front[0] = 0u;
    counts[0] = 1u;
    counts[1] = 0u;
    counts[2] = 0u;
    is_wide[0] = 1u;
}

// quantum exponent: smallest e (from 2^e >= extent / 255, then fixed up) with
// 255 * 2^e >= extent, all in double so the CPU build gets the same bits
__device__ int wide_exponent(double extent) {
    if (!(extent > 0.0)) return -100;
    int e;
    (void)frexp(extent / 255.0, &e);
    while (255.0 * ldexp(1.0, e) < extent) ++e;
    if (e < -100) e = -100;
    if (e > 100) e = 100;
    return e;
}

__global__ void k_wide_write(const float4* nodes, int nbin, const uint32_t* is_wide,
    const uint32_t* index, float4* nodes4) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbin || is_wide[b] == 0u) return;
    Wide w;
    expand_wide(nodes, (uint32_t)b, w);
    float org[3];
    double ext[3];
    for (int a = 0; a < 3; ++a) {
        float lo = __builtin_inff(), hi = -__builtin_inff();
        for (int k = 0; k < w.count; ++k) { lo = sel_min(lo, w.lo[k][a]); hi = sel_max(hi, w.hi[k][a]); }
        org[a] = lo;
        ext[a] = (double)hi - (double)lo;
    }
    uint32_t meta = (uint32_t)w.count << 24;
    uint32_t qlo[3] = {0u, 0u, 0u}, qhi[3] = {0u, 0u, 0u};
    for (int a = 0; a < 3; ++a) {
        const int e = wide_exponent(ext[a]);
        meta |= (uint32_t)(e + 127) << (8 * a);
        const double sc = ldexp(1.0, e), inv = ldexp(1.0, -e);
        for (int k = 0; k < w.count; ++k) {
            double ql = floor(((double)w.lo[k][a] - (double)org[a]) * inv);
            if (ql < 0.0) ql = 0.0;
            if (ql > 255.0) ql = 255.0;
            while (ql > 0.0 && (double)org[a] + ql * sc > (double)w.lo[k][a]) ql -= 1.0;
            double qh = ceil(((double)w.hi[k][a] - (double)org[a]) * inv);
            if (qh < 0.0) qh = 0.0;
            if (qh > 255.0) qh = 255.0;
            while (qh < 255.0 && (double)org[a] + qh * sc < (double)w.hi[k][a]) qh += 1.0;
            qlo[a] |= (uint32_t)ql << (8 * k);
            qhi[a] |= (uint32_t)qh << (8 * k);
        }
    }
    uint32_t ref[kWide];
    for (int k = 0; k < kWide; ++k) {
        if (k >= w.count) {
            ref[k] = kEmpty;
            continue;
        }
        const uint32_t nl = subtree_leaves(nodes, w.ref[k]);
        ref[k] = nl <= kLeafMax ? wide_leaf(first_leaf(nodes, w.ref[k]), nl) : index[w.ref[k]];
    }
    float4* out = nodes4 + 4u * index[b];
    out[0] = make_float4(org[0], org[1], org[2], __uint_as_float(meta));
    reinterpret_cast<uint4*>(out)[1] = make_uint4(ref[0], ref[1], ref[2], ref[3]);
    reinterpret_cast<uint4*>(out)[2] = make_uint4(qlo[0], qlo[1], qlo[2], qhi[0]);
    reinterpret_cast<uint4*>(out)[3] = make_uint4(qhi[1], qhi[2], 0u, 0u);
}

inline int blocks_for(int n, int b) { return (n + b - 1) / b; }

}  // namespace

hipError_t bvh_collapse_wide(const float4* nodes, int nbin, WideScratch& w, float4* nodes4,
    int* wide_count, int* wide_levels, hipStream_t st) {
    *wide_count = 0;
    *wide_levels = 0;
    if (nbin <= 0) return hipSuccess;
    hipError_t e = hipMemsetAsync(w.is_wide, 0, sizeof(uint32_t) * (size_t)nbin, st);
    if (e != hipSuccess) return e;
    k_wide_root<<<1, 64, 0, st>>>(w.front[0], w.counts, w.is_wide);
    // levels go out in batches with one host check per batch: a frontier
    // holds at most nbin nodes, and the grid-stride launch covers any size
    constexpr int kLevelBatch = 8;
    const int grid = blocks_for(nbin, 256) < 2048 ? blocks_for(nbin, 256) : 2048;
    int cur = 0;
    uint32_t host[3] = {1u, 0u, 0u};
    while (true) {
        for (int j = 0; j < kLevelBatch; ++j) {
            e = hipMemsetAsync(w.counts + (cur ^ 1), 0, sizeof(uint32_t), st);
            if (e != hipSuccess) return e;
            k_wide_level<<<grid, 256, 0, st>>>(nodes, w.front[cur], w.counts + cur, w.front[cur ^ 1],
                w.counts + (cur ^ 1), w.is_wide, w.counts + 2);
            cur ^= 1;
        }
        e = hipMemcpyAsync(host, w.counts, sizeof(host), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return e;
        if (host[cur] == 0u) break;
    }
    const int levels = (int)host[2];
    // compact numbering: exclusive scan of the marks
    const int tiles = scan_tiles(nbin);
    k_tile_sums<<<tiles, kScanBlock, 0, st>>>(w.is_wide, nbin, w.tiles);
    k_scan<<<1, 1024, 0, st>>>(w.tiles, tiles);
    k_tile_scan<<<tiles, kScanBlock, 0, st>>>(w.is_wide, nbin, w.tiles, w.index);
    uint32_t last[2] = {0u, 0u};
    e = hipMemcpyAsync(&last[0], w.index + (nbin - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(&last[1], w.is_wide + (nbin - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    *wide_count = (int)(last[0] + last[1]);
    *wide_levels = levels;
    k_wide_write<<<blocks_for(nbin, 256), 256, 0, st>>>(nodes, nbin, w.is_wide, w.index, nodes4);
    return hipGetLastError();
}

hipError_t bvh_build(const SceneView& s, int n, BvhBuildScratch& w, float4* nodes, float4* leaves,
    uint32_t* sorted_out, uint32_t* morton_out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    k_init_bounds<<<1, 64, 0, st>>>(w.cb);
    k_prim_bounds<<<bounds_blocks(n), kBoundsBlock, 0, st>>>(s, n, w.pbox, w.cb);
    k_morton<<<blocks_for(n, 256), 256, 0, st>>>(n, w.pbox, w.cb, w.keys[0], w.vals[0]);
    const int nblocks = blocks_for(n, kSortTile);
    int cur = 0;
    for (int shift = 0; shift < 32; shift += 4) {
        k_radix_count<<<nblocks, kSortBlock, 0, st>>>(w.keys[cur], n, shift, w.counts, nblocks);
        k_scan<<<1, 1024, 0, st>>>(w.counts, 16 * nblocks);
        k_radix_scatter<<<nblocks, kSortBlock, 0, st>>>(w.keys[cur], w.vals[cur], w.keys[cur ^ 1],
            w.vals[cur ^ 1], n, shift, w.counts, nblocks);
        cur ^= 1;
    }
    // 8 passes: the sorted pairs end in buffer 0
    hipError_t e = hipMemcpyAsync(sorted_out, w.vals[cur], sizeof(uint32_t) * (size_t)n,
        hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(morton_out, w.keys[cur], sizeof(uint32_t) * (size_t)n,
        hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return e;
    if (n == 1) {
        k_single<<<1, 64, 0, st>>>(s, sorted_out, w.pbox, nodes, leaves, w.cb);
        return hipGetLastError();
    }
    e = hipMemsetAsync(w.flags, 0, sizeof(uint32_t) * (size_t)(n - 1), st);
    if (e != hipSuccess) return e;
    k_karras<<<blocks_for(n - 1, 256), 256, 0, st>>>(morton_out, n,
        reinterpret_cast<uint4*>(nodes), w.leaf_parent);
    k_refit<<<blocks_for(n, 256), 256, 0, st>>>(s, n, sorted_out, w.pbox, w.leaf_parent, nodes,
        w.flags, leaves);
    k_depth<<<blocks_for(n, 256), 256, 0, st>>>(n, w.leaf_parent, nodes, w.cb);
    return hipGetLastError();
}

hipError_t bvh_refit(const SceneView& s, int n, BvhBuildScratch& w, float4* nodes, float4* leaves,
    const uint32_t* sorted, const uint32_t* leaf_parent, hipStream_t st) {
    if (n <= 1) return hipErrorInvalidValue;
    k_init_bounds<<<1, 64, 0, st>>>(w.cb);
    k_prim_bounds<<<bounds_blocks(n), kBoundsBlock, 0, st>>>(s, n, w.pbox, w.cb);
    const hipError_t e = hipMemsetAsync(w.flags, 0, sizeof(uint32_t) * (size_t)(n - 1), st);
    if (e != hipSuccess) return e;
    // the build's bottom-up pass over the kept topology: new leaf records,
    // new child boxes (exact min / max, so order-free and bit-identical to a
    // sequential refit)
    k_refit<<<blocks_for(n, 256), 256, 0, st>>>(s, n, sorted, w.pbox, leaf_parent, nodes, w.flags,
        leaves);
    return hipGetLastError();
}

}  // namespace rtk
