/*
 * oracle_bvh.c -- CPU Morton LBVH, traversal, brute force, intersectors.
 * TEST INFRASTRUCTURE ONLY (see rt_oracle.h).
 *
 * Build (the reference builds a median-split BVH on the CPU,
 * src/realtime/gpu/gpu_scene_acceleration.cpp:145-179; the target is a
 * Morton LBVH, so this restates Karras 2012, "Maximizing Parallelism in the
 * Construction of BVHs, Octrees, and k-d Trees", HPG 2012, Sec. 4):
 *   1. primitive bounds: sphere c +- |r|; quad = its 4 corners, triangle = its
 *      3 vertices, both padded by 1e-5 on axes thinner than 1e-5
 *      (gpu_scene_acceleration.cpp:19,43-71);
 *   2. centroid = 0.5*(lo+hi); quantised to 10 bits per axis inside the
 *      centroid bounds; 30-bit Morton code x:y:z interleaved (x highest);
 *   3. stable sort by code (ties keep primitive-id order);
 *   4. Karras radix tree; delta(i,j) = clz(k_i ^ k_j), or 32 + clz(i ^ j) for
 *      equal codes; out-of-range j gives -1;
 *   5. exact min/max refit; each node stores its two children's boxes.
 * Primitive ids are global: spheres, then quads, then triangles.
 *
 * Closest hit: the winning primitive is the lexicographic minimum of
 * (t, primitive id) over all primitives with tmin < t < tmax, so BVH
 * traversal and brute force must agree exactly.  Intersectors follow
 * programs.cu:1421-1547 operation for operation.
 */
#include "rt_oracle.h"
#include "om.h"

#include <float.h>
#include <pthread.h>
#include <stdlib.h>

int rto_version(void) { return 1; }

void rto_pcg4d(uint32_t v[4]) { om_pcg4d(v); }
void rto_rng4(uint32_t pixel, uint32_t sample, uint32_t camera, uint32_t bounce, uint32_t group,
    uint32_t stream, float out[4]) {
    om_rng4(pixel, sample, camera, bounce, group, stream, out);
}
void rto_sincos_turn(float u, float* s, float* c) { om_sincos_turn(u, s, c); }
float rto_cbrt01(float x) { return om_cbrt01(x); }

/* ---------------------------------------------------------------- bounds */
typedef struct { float lo[3]; float hi[3]; } box3;

static void box_empty(box3* b) {
    for (int a = 0; a < 3; ++a) { b->lo[a] = INFINITY; b->hi[a] = -INFINITY; }
}
static void box_point(box3* b, v3 p) {
    const float c[3] = {p.x, p.y, p.z};
    for (int a = 0; a < 3; ++a) { b->lo[a] = sel_min(b->lo[a], c[a]); b->hi[a] = sel_max(b->hi[a], c[a]); }
}
static void box_union(box3* b, const box3* o) {
    for (int a = 0; a < 3; ++a) { b->lo[a] = sel_min(b->lo[a], o->lo[a]); b->hi[a] = sel_max(b->hi[a], o->hi[a]); }
}
static void box_pad(box3* b) {
    for (int a = 0; a < 3; ++a) {
        if (b->hi[a] - b->lo[a] < 1e-5f) { b->lo[a] -= 1e-5f; b->hi[a] += 1e-5f; }
    }
}

static int prim_total(const rto_scene* s) { return s->sphere_count + s->quad_count + s->triangle_count; }

static void prim_bounds(const rto_scene* s, int id, box3* b) {
    box_empty(b);
    if (id < s->sphere_count) {
        const rto_sphere* sp = &s->spheres[id];
        const float r = fabsf(sp->radius);
        for (int a = 0; a < 3; ++a) { b->lo[a] = sp->center[a] - r; b->hi[a] = sp->center[a] + r; }
        return;
    }
    id -= s->sphere_count;
    if (id < s->quad_count) {
        const rto_quad* q = &s->quads[id];
        const v3 o = ld3(q->origin), u = ld3(q->edge_u), v = ld3(q->edge_v);
        box_point(b, o);
        box_point(b, add3(o, u));
        box_point(b, add3(o, v));
        box_point(b, add3(add3(o, u), v));
        box_pad(b);
        return;
    }
    id -= s->quad_count;
    const rto_triangle* t = &s->triangles[id];
    box_point(b, ld3(t->p0));
    box_point(b, ld3(t->p1));
    box_point(b, ld3(t->p2));
    box_pad(b);
}

uint32_t rto_expand_bits10(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

static uint32_t quantize10(float c, float lo, float scale) {
    float q = (c - lo) * scale;
    q = q < 0.0f ? 0.0f : q;
    q = q > 1023.0f ? 1023.0f : q;
    return (uint32_t)q;
}

static int clz32(uint32_t x) { return x == 0 ? 32 : __builtin_clz(x); }

static int delta_fn(const uint32_t* keys, int n, int i, int j) {
    if (j < 0 || j > n - 1) return -1;
    const uint32_t a = keys[i], b = keys[j];
    if (a == b) return 32 + clz32((uint32_t)i ^ (uint32_t)j);
    return clz32(a ^ b);
}

/* stable LSD radix sort of (key, value) pairs, 8 bits per pass */
static void radix_sort_pairs(uint32_t* keys, uint32_t* vals, int n) {
    uint32_t* k2 = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    uint32_t* v2 = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    for (int shift = 0; shift < 32; shift += 8) {
        size_t count[257] = {0};
        for (int i = 0; i < n; ++i) count[((keys[i] >> shift) & 0xFFu) + 1]++;
        for (int b = 0; b < 256; ++b) count[b + 1] += count[b];
        for (int i = 0; i < n; ++i) {
            const size_t dst = count[(keys[i] >> shift) & 0xFFu]++;
            k2[dst] = keys[i];
            v2[dst] = vals[i];
        }
        memcpy(keys, k2, sizeof(uint32_t) * (size_t)n);
        memcpy(vals, v2, sizeof(uint32_t) * (size_t)n);
    }
    free(k2);
    free(v2);
}

static void node_set_child_box(rto_bvh_node* nd, int slot, const box3* b) {
    float* lo = slot == 0 ? nd->lo0 : nd->lo1;
    float* hi = slot == 0 ? nd->hi0 : nd->hi1;
    for (int a = 0; a < 3; ++a) { lo[a] = b->lo[a]; hi[a] = b->hi[a]; }
}

static int collapse_wide(rto_bvh* bvh);

int rto_build_lbvh(const rto_scene* s, rto_bvh* out) {
    memset(out, 0, sizeof(*out));
    const int n = prim_total(s);
    out->prim_count = n;
    if (n <= 0) return 0;
    box3* pb = (box3*)malloc(sizeof(box3) * (size_t)n);
    box3 cb;
    box_empty(&cb);
    for (int i = 0; i < n; ++i) {
        prim_bounds(s, i, &pb[i]);
        const v3 c = mk3((pb[i].lo[0] + pb[i].hi[0]) * 0.5f, (pb[i].lo[1] + pb[i].hi[1]) * 0.5f,
            (pb[i].lo[2] + pb[i].hi[2]) * 0.5f);
        box_point(&cb, c);
    }
    float scale[3];
    for (int a = 0; a < 3; ++a) {
        const float ext = cb.hi[a] - cb.lo[a];
        scale[a] = ext > 0.0f ? 1024.0f / ext : 0.0f;
    }
    uint32_t* keys = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
    uint32_t* vals = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
    for (int i = 0; i < n; ++i) {
        const float c0 = (pb[i].lo[0] + pb[i].hi[0]) * 0.5f;
        const float c1 = (pb[i].lo[1] + pb[i].hi[1]) * 0.5f;
        const float c2 = (pb[i].lo[2] + pb[i].hi[2]) * 0.5f;
        const uint32_t qx = quantize10(c0, cb.lo[0], scale[0]);
        const uint32_t qy = quantize10(c1, cb.lo[1], scale[1]);
        const uint32_t qz = quantize10(c2, cb.lo[2], scale[2]);
        keys[i] = (rto_expand_bits10(qx) << 2) | (rto_expand_bits10(qy) << 1) | rto_expand_bits10(qz);
        vals[i] = (uint32_t)i;
    }
    radix_sort_pairs(keys, vals, n);

    const int nodes = n > 1 ? n - 1 : 1;
    out->node_count = nodes;
    out->nodes = (rto_bvh_node*)calloc((size_t)nodes, sizeof(rto_bvh_node));
    out->sorted_prims = vals;
    out->morton = keys;
    uint32_t* leaf_parent = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);

    if (n == 1) {
        rto_bvh_node* r = &out->nodes[0];
        r->child[0] = OM_LEAF_BIT | 0u;
        r->child[1] = OM_EMPTY;
        r->parent = OM_EMPTY;
        r->pad = 1u;
        node_set_child_box(r, 0, &pb[vals[0]]);
        box3 e;
        box_empty(&e);
        node_set_child_box(r, 1, &e);
        out->max_depth = 1;
        free(leaf_parent);
        free(pb);
        return collapse_wide(out);
    }

    out->nodes[0].parent = OM_EMPTY;
    for (int i = 0; i < n - 1; ++i) {
        const int d = (delta_fn(keys, n, i, i + 1) - delta_fn(keys, n, i, i - 1)) >= 0 ? 1 : -1;
        const int dmin = delta_fn(keys, n, i, i - d);
        int lmax = 2;
        while (delta_fn(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
        int l = 0;
        for (int t = lmax / 2; t >= 1; t /= 2) {
            if (delta_fn(keys, n, i, i + (l + t) * d) > dmin) l += t;
        }
        const int j = i + l * d;
        const int dnode = delta_fn(keys, n, i, j);
        int s_ = 0;
        int div = 2;
        int t = (l + div - 1) / div;
        while (1) {
            if (delta_fn(keys, n, i, i + (s_ + t) * d) > dnode) s_ += t;
            if (t == 1) break;
            div *= 2;
            t = (l + div - 1) / div;
        }
        const int gamma = i + s_ * d + (d < 0 ? d : 0);
        const int first = i < j ? i : j;
        const int last = i < j ? j : i;
        rto_bvh_node* nd = &out->nodes[i];
        nd->pad = (uint32_t)(last - first + 1);  /* leaves under this node */
        if (first == gamma) {
            nd->child[0] = OM_LEAF_BIT | (uint32_t)gamma;
            leaf_parent[gamma] = (uint32_t)i;
        } else {
            nd->child[0] = (uint32_t)gamma;
            out->nodes[gamma].parent = (uint32_t)i;
        }
        if (last == gamma + 1) {
            nd->child[1] = OM_LEAF_BIT | (uint32_t)(gamma + 1);
            leaf_parent[gamma + 1] = (uint32_t)i;
        } else {
            nd->child[1] = (uint32_t)(gamma + 1);
            out->nodes[gamma + 1].parent = (uint32_t)i;
        }
    }
    /* refit: post-order by explicit stack (exact min/max, order-free) */
    box3* nb = (box3*)malloc(sizeof(box3) * (size_t)nodes);
    int* stack = (int*)malloc(sizeof(int) * (size_t)(2 * nodes + 2));
    unsigned char* done = (unsigned char*)calloc((size_t)nodes, 1);
    int sp = 0;
    stack[sp++] = 0;
    int max_depth = 0;
    int* depth = (int*)calloc((size_t)nodes, sizeof(int));
    depth[0] = 1;
    while (sp > 0) {
        const int i = stack[sp - 1];
        rto_bvh_node* nd = &out->nodes[i];
        if (!done[i]) {
            done[i] = 1;
            for (int c = 0; c < 2; ++c) {
                if (!(nd->child[c] & OM_LEAF_BIT)) {
                    depth[nd->child[c]] = depth[i] + 1;
                    stack[sp++] = (int)nd->child[c];
                }
            }
            if (depth[i] > max_depth) max_depth = depth[i];
            continue;
        }
        --sp;
        box3 acc;
        box_empty(&acc);
        for (int c = 0; c < 2; ++c) {
            const uint32_t ch = nd->child[c];
            const box3* b = (ch & OM_LEAF_BIT) ? &pb[vals[ch & ~OM_LEAF_BIT]] : &nb[ch];
            node_set_child_box(nd, c, b);
            box_union(&acc, b);
        }
        nb[i] = acc;
    }
    out->max_depth = max_depth;
    if (collapse_wide(out) != 0) { free(depth); free(done); free(stack); free(nb); free(leaf_parent); free(pb); return -1; }
    free(depth);
    free(done);
    free(stack);
    free(nb);
    free(leaf_parent);
    free(pb);
    return 0;
}

/* Refit (GpuSceneAcceleration::refit, gpu_scene_acceleration.cpp:348-367,
 * over this LBVH): the binary topology and the Morton-sorted leaf order of the
 * last build are kept; every child box is recomputed bottom-up from the
 * primitive bounds of `s` (the same padded bounds and exact min / max as the
 * build), then the 4-wide tree is collapsed again.  s must have the same
 * primitive counts as the scene the BVH was built for. */
int rto_refit_lbvh(const rto_scene* s, rto_bvh* bvh) {
    const int n = prim_total(s);
    if (n != bvh->prim_count || n <= 1) return -1;
    box3* pb = (box3*)malloc(sizeof(box3) * (size_t)n);
    for (int i = 0; i < n; ++i) prim_bounds(s, i, &pb[i]);
    const int nodes = bvh->node_count;
    box3* nb = (box3*)malloc(sizeof(box3) * (size_t)nodes);
    int* stack = (int*)malloc(sizeof(int) * (size_t)(2 * nodes + 2));
    unsigned char* done = (unsigned char*)calloc((size_t)nodes, 1);
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const int i = stack[sp - 1];
        rto_bvh_node* nd = &bvh->nodes[i];
        if (!done[i]) {
            done[i] = 1;
            for (int c = 0; c < 2; ++c)
                if (!(nd->child[c] & OM_LEAF_BIT)) stack[sp++] = (int)nd->child[c];
            continue;
        }
        --sp;
        box3 acc;
        box_empty(&acc);
        for (int c = 0; c < 2; ++c) {
            const uint32_t ch = nd->child[c];
            const box3* b = (ch & OM_LEAF_BIT) ? &pb[bvh->sorted_prims[ch & ~OM_LEAF_BIT]] : &nb[ch];
            node_set_child_box(nd, c, b);
            box_union(&acc, b);
        }
        nb[i] = acc;
    }
    free(done);
    free(stack);
    free(nb);
    free(pb);
    free(bvh->wide);
    bvh->wide = NULL;
    bvh->wide_count = 0;
    return collapse_wide(bvh);
}

/* ------------------------------------------------- 4-wide collapse
 * expand(b): start from binary node b's two children; while fewer than 4,
 * replace the internal child with the largest surface area (first on ties)
 * by its two children, in place.  Nodes are numbered by an exclusive scan of
 * the "roots a 4-wide node" marks in binary node order. */
typedef struct { uint32_t ref[4]; float lo[4][3]; float hi[4][3]; int count; } wide_t;

static float box_area3(const float* lo, const float* hi) {
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return dx * dy + dy * dz + dz * dx;
}

static void binary_children(const rto_bvh_node* nodes, uint32_t b, uint32_t ref[2], float lo[2][3],
    float hi[2][3]) {
    const rto_bvh_node* nd = &nodes[b];
    ref[0] = nd->child[0];
    ref[1] = nd->child[1];
    for (int a = 0; a < 3; ++a) {
        lo[0][a] = nd->lo0[a]; hi[0][a] = nd->hi0[a];
        lo[1][a] = nd->lo1[a]; hi[1][a] = nd->hi1[a];
    }
}

#define WIDE_LEAF_MAX 2u
static uint32_t subtree_leaves(const rto_bvh_node* nodes, uint32_t ref) {
    return (ref & OM_LEAF_BIT) ? 1u : nodes[ref].pad;
}
static uint32_t first_leaf(const rto_bvh_node* nodes, uint32_t ref) {
    while (!(ref & OM_LEAF_BIT)) ref = nodes[ref].child[0];
    return ref & ~OM_LEAF_BIT;
}

static void expand_wide(const rto_bvh_node* nodes, uint32_t b, wide_t* w) {
    uint32_t r2[2];
    float l2[2][3], h2[2][3];
    binary_children(nodes, b, r2, l2, h2);
    w->count = 0;
    for (int c = 0; c < 2; ++c) {
        if (r2[c] == OM_EMPTY) continue;
        w->ref[w->count] = r2[c];
        for (int a = 0; a < 3; ++a) { w->lo[w->count][a] = l2[c][a]; w->hi[w->count][a] = h2[c][a]; }
        ++w->count;
    }
    while (w->count < 4) {
        int best = -1;
        float best_area = -1.0f;
        for (int j = 0; j < w->count; ++j) {
            if (subtree_leaves(nodes, w->ref[j]) <= WIDE_LEAF_MAX) continue;
            const float ar = box_area3(w->lo[j], w->hi[j]);
            if (ar > best_area) { best_area = ar; best = j; }
        }
        if (best < 0) break;
        binary_children(nodes, w->ref[best], r2, l2, h2);
        for (int j = w->count; j > best + 1; --j) {
            w->ref[j] = w->ref[j - 1];
            for (int a = 0; a < 3; ++a) { w->lo[j][a] = w->lo[j - 1][a]; w->hi[j][a] = w->hi[j - 1][a]; }
        }
        for (int c = 0; c < 2; ++c) {
            w->ref[best + c] = r2[c];
            for (int a = 0; a < 3; ++a) { w->lo[best + c][a] = l2[c][a]; w->hi[best + c][a] = h2[c][a]; }
        }
        ++w->count;
    }
}

/* smallest quantum 2^e (from frexp, then fixed up) with 255 * 2^e >= extent */
static int wide_exponent(double extent) {
    if (!(extent > 0.0)) return -100;
    int e;
    (void)frexp(extent / 255.0, &e);
    while (255.0 * ldexp(1.0, e) < extent) ++e;
    if (e < -100) e = -100;
    if (e > 100) e = 100;
    return e;
}

static int collapse_wide(rto_bvh* bvh) {
    const int nbin = bvh->node_count;
    bvh->wide = NULL;
    bvh->wide_count = 0;
    bvh->wide_levels = 0;
    if (nbin <= 0) return 0;
    uint32_t* is_wide = (uint32_t*)calloc((size_t)nbin, sizeof(uint32_t));
    uint32_t* front = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)nbin);
    uint32_t* next = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)nbin);
    uint32_t* index = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)nbin);
    if (!is_wide || !front || !next || !index) { free(is_wide); free(front); free(next); free(index); return -1; }
    is_wide[0] = 1u;
    front[0] = 0u;
    int nf = 1, levels = 0;
    while (nf > 0) {
        ++levels;
        int nn = 0;
        for (int i = 0; i < nf; ++i) {
            wide_t w;
            expand_wide(bvh->nodes, front[i], &w);
            for (int k = 0; k < w.count; ++k) {
                if (subtree_leaves(bvh->nodes, w.ref[k]) <= WIDE_LEAF_MAX) continue;
                is_wide[w.ref[k]] = 1u;
                next[nn++] = w.ref[k];
            }
        }
        uint32_t* t = front; front = next; next = t;
        nf = nn;
    }
    uint32_t run = 0;
    for (int b = 0; b < nbin; ++b) { index[b] = run; run += is_wide[b]; }
    rto_wide_node* out = (rto_wide_node*)calloc(run, sizeof(rto_wide_node));
    if (!out) { free(is_wide); free(front); free(next); free(index); return -1; }
    for (int b = 0; b < nbin; ++b) {
        if (!is_wide[b]) continue;
        wide_t w;
        expand_wide(bvh->nodes, (uint32_t)b, &w);
        rto_wide_node* o = &out[index[b]];
        double ext[3];
        for (int a = 0; a < 3; ++a) {
            float lo = INFINITY, hi = -INFINITY;
            for (int k = 0; k < w.count; ++k) { lo = sel_min(lo, w.lo[k][a]); hi = sel_max(hi, w.hi[k][a]); }
            o->origin[a] = lo;
            ext[a] = (double)hi - (double)lo;
        }
        uint32_t meta = (uint32_t)w.count << 24;
        for (int a = 0; a < 3; ++a) {
            const int e = wide_exponent(ext[a]);
            meta |= (uint32_t)(e + 127) << (8 * a);
            const double sc = ldexp(1.0, e), inv = ldexp(1.0, -e);
            for (int k = 0; k < w.count; ++k) {
                double ql = floor(((double)w.lo[k][a] - (double)o->origin[a]) * inv);
                if (ql < 0.0) ql = 0.0;
                if (ql > 255.0) ql = 255.0;
                while (ql > 0.0 && (double)o->origin[a] + ql * sc > (double)w.lo[k][a]) ql -= 1.0;
                double qh = ceil(((double)w.hi[k][a] - (double)o->origin[a]) * inv);
                if (qh < 0.0) qh = 0.0;
                if (qh > 255.0) qh = 255.0;
                while (qh < 255.0 && (double)o->origin[a] + qh * sc < (double)w.hi[k][a]) qh += 1.0;
                o->qlo[a] |= (uint32_t)ql << (8 * k);
                o->qhi[a] |= (uint32_t)qh << (8 * k);
            }
        }
        o->meta = meta;
        for (int k = 0; k < 4; ++k) {
            if (k >= w.count) { o->child[k] = OM_EMPTY; continue; }
            const uint32_t nl = subtree_leaves(bvh->nodes, w.ref[k]);
            o->child[k] = nl <= WIDE_LEAF_MAX
                              ? (OM_LEAF_BIT | ((nl - 1u) << 28) | first_leaf(bvh->nodes, w.ref[k]))
                              : index[w.ref[k]];
        }
    }
    bvh->wide = out;
    bvh->wide_count = (int32_t)run;
    bvh->wide_levels = levels;
    free(is_wide); free(front); free(next); free(index);
    return 0;
}

void rto_free_lbvh(rto_bvh* bvh) {
    free(bvh->wide);
    free(bvh->nodes);
    free(bvh->sorted_prims);
    free(bvh->morton);
    memset(bvh, 0, sizeof(*bvh));
}

/* structural check: every leaf reached exactly once, boxes contain children */
int rto_bvh_validate(const rto_scene* s, const rto_bvh* bvh) {
    const int n = bvh->prim_count;
    if (n == 0) return 0;
    unsigned char* seen = (unsigned char*)calloc((size_t)n, 1);
    int* stack = (int*)malloc(sizeof(int) * (size_t)(bvh->node_count + 2));
    int sp = 0, err = 0;
    stack[sp++] = 0;
    while (sp > 0 && !err) {
        const rto_bvh_node* nd = &bvh->nodes[stack[--sp]];
        for (int c = 0; c < 2; ++c) {
            const uint32_t ch = nd->child[c];
            if (ch == OM_EMPTY) continue;
            if (ch & OM_LEAF_BIT) {
                const uint32_t leaf = ch & ~OM_LEAF_BIT;
                if ((int)leaf >= n || seen[leaf]) { err = 1; break; }
                seen[leaf] = 1;
                box3 b;
                prim_bounds(s, (int)bvh->sorted_prims[leaf], &b);
                const float* lo = c == 0 ? nd->lo0 : nd->lo1;
                const float* hi = c == 0 ? nd->hi0 : nd->hi1;
                for (int a = 0; a < 3; ++a)
                    if (lo[a] != b.lo[a] || hi[a] != b.hi[a]) err = 2;
            } else {
                if ((int)ch >= bvh->node_count) { err = 3; break; }
                stack[sp++] = (int)ch;
            }
        }
    }
    for (int i = 0; i < n && !err; ++i)
        if (!seen[i]) err = 4;
    free(stack);
    free(seen);
    return err;
}

/* ----------------------------------------------------------- intersectors */
/* programs.cu:1421-1449; returns t in (tmin, tmax) or -1 */
static float isect_sphere(const rto_sphere* sp, v3 o, v3 d, float tmin, float tmax) {
    const v3 oc = sub3(o, ld3(sp->center));
    const float a = lensq3(d);
    const float half_b = dot3(oc, d);
    const float c = lensq3(oc) - sp->radius * sp->radius;
    const float disc = half_b * half_b - a * c;
    if (disc <= 0.0f) return -1.0f;
    const float sq = sqrtf(disc);
    float root = (-half_b - sq) / a;
    if (root <= tmin || root >= tmax) {
        root = (-half_b + sq) / a;
        if (root <= tmin || root >= tmax) return -1.0f;
    }
    return root;
}

/* programs.cu:1451-1488 */
static float isect_quad(const rto_quad* q, v3 o, v3 d, float tmin, float tmax, float* pu, float* pv) {
    const v3 qo = ld3(q->origin), eu = ld3(q->edge_u), ev = ld3(q->edge_v);
    const v3 n = cross3(eu, ev);
    const float nls = lensq3(n);
    if (nls <= 1e-12f) return -1.0f;
    const v3 normal = div3(n, sqrtf(nls));
    const float denom = dot3(normal, d);
    if (fabsf(denom) < 1e-6f) return -1.0f;
    const float t = dot3(sub3(qo, o), normal) / denom;
    if (t <= tmin || t >= tmax) return -1.0f;
    const v3 p = add3(o, scl3(d, t));
    const v3 rel = sub3(p, qo);
    const v3 w = div3(n, nls);
    const float u = dot3(w, cross3(rel, ev));
    const float v = dot3(w, cross3(eu, rel));
    if (u < 0.0f || u > 1.0f || v < 0.0f || v > 1.0f) return -1.0f;
    *pu = u;
    *pv = v;
    return t;
}

/* programs.cu:1490-1520 (Moller-Trumbore) */
static float isect_tri(const rto_triangle* tr, v3 o, v3 d, float tmin, float tmax, float* pu, float* pv) {
    const v3 p0 = ld3(tr->p0);
    const v3 e1 = sub3(ld3(tr->p1), p0);
    const v3 e2 = sub3(ld3(tr->p2), p0);
    const v3 pvec = cross3(d, e2);
    const float det = dot3(e1, pvec);
    if (fabsf(det) <= 1e-8f) return -1.0f;
    const float inv_det = 1.0f / det;
    const v3 tvec = sub3(o, p0);
    const float u = dot3(tvec, pvec) * inv_det;
    if (u < 0.0f || u > 1.0f) return -1.0f;
    const v3 qvec = cross3(tvec, e1);
    const float v = dot3(d, qvec) * inv_det;
    if (v < 0.0f || (u + v) > 1.0f) return -1.0f;
    const float t = dot3(e2, qvec) * inv_det;
    if (t <= tmin || t >= tmax) return -1.0f;
    *pu = u;
    *pv = v;
    return t;
}

static float isect_prim(const rto_scene* s, int id, v3 o, v3 d, float tmin, float tmax, float* u, float* v) {
    *u = 0.0f;
    *v = 0.0f;
    if (id < s->sphere_count) return isect_sphere(&s->spheres[id], o, d, tmin, tmax);
    id -= s->sphere_count;
    if (id < s->quad_count) return isect_quad(&s->quads[id], o, d, tmin, tmax, u, v);
    id -= s->quad_count;
    return isect_tri(&s->triangles[id], o, d, tmin, tmax, u, v);
}

/* ---------------------------------------------------------------- traversal */
typedef struct { float t; int32_t prim; float u; float v; int found; } best_hit;

static void consider(best_hit* b, int id, float t, float u, float v) {
    if (t < 0.0f) return;
    if (t < b->t || (t == b->t && b->found && id < b->prim)) {
        b->t = t; b->prim = id; b->u = u; b->v = v; b->found = 1;
    }
}

/* Traversal of the 4-wide quantized BVH (shared definition for the
 * box-test / prim-test counts, which the GPU kernels reproduce):
 *  - at a node every valid child box is tested against the current best
 *    distance (one counted box test each);
 *  - leaf-range children that pass are intersected in slot order, their
 *    primitives in sorted order;
 *  - internal children that pass are ordered by entry distance (stable: the
 *    lower slot first on ties); the nearest is visited next and the others
 *    are pushed farthest first with their entry distance;
 *  - popped entries whose entry distance exceeds the best distance are
 *    skipped.
 * Quantized box test: per axis t = fma(q, inv * 2^e, (origin - o) * inv);
 * the exit distance is widened by 2.4e-6 relative plus 4e-7 of the summed
 * |(origin - o) * inv| so rounding never culls a hit on a box face. */
typedef struct { float inv[3]; } ray_pre;

static int wide_box_tests(const rto_wide_node* nd, const float* o, const ray_pre* p, float tmin,
    float tbest, float tn_out[4]) {
    float sc[3], off[3];
    for (int a = 0; a < 3; ++a) {
        uint32_t eb = (nd->meta >> (8 * a)) & 255u;
        uint32_t bits = eb << 23;
        float pw;
        memcpy(&pw, &bits, 4);
        sc[a] = p->inv[a] * pw;
        off[a] = (nd->origin[a] - o[a]) * p->inv[a];
    }
    const float slack = 4e-7f * (fabsf(off[0]) + fabsf(off[1]) + fabsf(off[2]));
    int mask = 0;
    for (int k = 0; k < 4; ++k) {
        tn_out[k] = 0.0f;
        if (nd->child[k] == OM_EMPTY) continue;
        float t0[3], t1[3];
        for (int a = 0; a < 3; ++a) {
            const float ql = (float)((nd->qlo[a] >> (8 * k)) & 255u);
            const float qh = (float)((nd->qhi[a] >> (8 * k)) & 255u);
            t0[a] = fmaf(ql, sc[a], off[a]);
            t1[a] = fmaf(qh, sc[a], off[a]);
        }
        const float nx = fminf(t0[0], t1[0]), ny = fminf(t0[1], t1[1]), nz = fminf(t0[2], t1[2]);
        const float fx = fmaxf(t0[0], t1[0]), fy = fmaxf(t0[1], t1[1]), fz = fmaxf(t0[2], t1[2]);
        const float tn = fmaxf(fmaxf(nx, ny), fmaxf(nz, tmin));
        const float tf = fminf(fminf(fx, fy), fminf(fz, tbest)) * 1.0000024f + slack;
        tn_out[k] = tn;
        if (tn <= tf) mask |= 1 << k;
    }
    return mask;
}

static void traverse(const rto_scene* s, const rto_bvh* bvh, const rto_ray* r, int any,
    best_hit* out, uint32_t* visits, uint32_t* tests) {
    best_hit b = {r->tmax, -1, 0.0f, 0.0f, 0};
    uint32_t nv = 0, nt = 0;
    if (bvh->prim_count > 0 && bvh->wide_count > 0) {
        const v3 o = ld3(r->o), d = ld3(r->d);
        ray_pre p;
        const float dd[3] = {d.x, d.y, d.z};
        for (int a = 0; a < 3; ++a) {
            const float dv = fabsf(dd[a]) > 1e-12f ? dd[a] : copysignf(1e-12f, dd[a]);
            p.inv[a] = 1.0f / dv;
        }
        uint32_t stack_node[192];
        float stack_t[192];
        int sp = 0;
        uint32_t node = 0;
        while (1) {
            const rto_wide_node* nd = &bvh->wide[node];
            nv += nd->meta >> 24;
            float tn[4];
            const int mask = wide_box_tests(nd, r->o, &p, r->tmin, b.t, tn);
            for (int k = 0; k < 4; ++k) {
                if (!(mask & (1 << k)) || !(nd->child[k] & OM_LEAF_BIT)) continue;
                const uint32_t first = nd->child[k] & 0x0FFFFFFFu;
                const uint32_t cnt = ((nd->child[k] >> 28) & 7u) + 1u;
                for (uint32_t i = 0; i < cnt; ++i) {
                    const int id = (int)bvh->sorted_prims[first + i];
                    float u, v;
                    ++nt;
                    const float t = isect_prim(s, id, o, d, r->tmin, r->tmax, &u, &v);
                    consider(&b, id, t, u, v);
                    if (any && b.found) goto done;
                }
            }
            float key[4];
            uint32_t ref[4];
            int m = 0;
            for (int k = 0; k < 4; ++k) {
                if (!(mask & (1 << k)) || (nd->child[k] & OM_LEAF_BIT)) continue;
                int j = m++;
                while (j > 0 && tn[k] < key[j - 1]) { key[j] = key[j - 1]; ref[j] = ref[j - 1]; --j; }
                key[j] = tn[k];
                ref[j] = nd->child[k];
            }
            if (m > 0) {
                for (int j = m - 1; j >= 1; --j) { stack_node[sp] = ref[j]; stack_t[sp] = key[j]; ++sp; }
                node = ref[0];
                continue;
            }
            int got = 0;
            while (sp > 0) {
                --sp;
                if (stack_t[sp] <= b.t) { node = stack_node[sp]; got = 1; break; }
            }
            if (!got) break;
        }
    }
done:
    *out = b;
    if (visits) *visits = nv;
    if (tests) *tests = nt;
}

typedef struct {
    const rto_scene* s; const rto_bvh* bvh; const rto_ray* rays; int n; int any;
    rto_hit* hits; int32_t* occ; uint32_t* visits; uint32_t* tests; int tid; int nthreads;
} trace_job;

static void* trace_worker(void* arg) {
    trace_job* j = (trace_job*)arg;
    for (int i = j->tid; i < j->n; i += j->nthreads) {
        best_hit b;
        uint32_t nv, nt;
        traverse(j->s, j->bvh, &j->rays[i], j->any, &b, &nv, &nt);
        if (j->visits) j->visits[i] = nv;
        if (j->tests) j->tests[i] = nt;
        if (j->any) {
            j->occ[i] = b.found ? 1 : 0;
        } else {
            j->hits[i].t = b.found ? b.t : -1.0f;
            j->hits[i].prim = b.found ? b.prim : -1;
            j->hits[i].u = b.u;
            j->hits[i].v = b.v;
        }
    }
    return NULL;
}

static void run_trace(trace_job* base, int threads) {
    if (threads < 1) threads = 1;
    if (threads > 64) threads = 64;
    pthread_t th[64];
    trace_job jobs[64];
    for (int t = 0; t < threads; ++t) {
        jobs[t] = *base;
        jobs[t].tid = t;
        jobs[t].nthreads = threads;
        pthread_create(&th[t], NULL, trace_worker, &jobs[t]);
    }
    for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
}

void rto_trace_closest(const rto_scene* s, const rto_bvh* bvh, const rto_ray* rays, int n,
    rto_hit* hits, uint32_t* visits, uint32_t* tests, int threads) {
    trace_job j = {s, bvh, rays, n, 0, hits, NULL, visits, tests, 0, 1};
    run_trace(&j, threads);
}

void rto_trace_any(const rto_scene* s, const rto_bvh* bvh, const rto_ray* rays, int n,
    int32_t* occluded, uint32_t* visits, uint32_t* tests, int threads) {
    trace_job j = {s, bvh, rays, n, 1, NULL, occluded, visits, tests, 0, 1};
    run_trace(&j, threads);
}

void rto_brute_closest(const rto_scene* s, const rto_ray* rays, int n, rto_hit* hits) {
    const int np = prim_total(s);
    for (int i = 0; i < n; ++i) {
        const rto_ray* r = &rays[i];
        best_hit b = {r->tmax, -1, 0.0f, 0.0f, 0};
        const v3 o = ld3(r->o), d = ld3(r->d);
        for (int id = 0; id < np; ++id) {
            float u, v;
            const float t = isect_prim(s, id, o, d, r->tmin, r->tmax, &u, &v);
            consider(&b, id, t, u, v);
        }
        hits[i].t = b.found ? b.t : -1.0f;
        hits[i].prim = b.found ? b.prim : -1;
        hits[i].u = b.u;
        hits[i].v = b.v;
    }
}

/* internal entries used by the renderer; the counted form adds the
 * traversal's child-box and primitive tests to *visits / *tests */
int om_closest_counted(const rto_scene* s, const rto_bvh* bvh, v3 o, v3 d, float tmin, float tmax,
    float* t, int* prim, float* u, float* v, uint64_t* visits, uint64_t* tests) {
    rto_ray r = {{o.x, o.y, o.z}, tmin, {d.x, d.y, d.z}, tmax};
    best_hit b;
    uint32_t nv = 0, nt = 0;
    traverse(s, bvh, &r, 0, &b, &nv, &nt);
    if (visits) *visits += nv;
    if (tests) *tests += nt;
    *t = b.t; *prim = b.prim; *u = b.u; *v = b.v;
    return b.found;
}

int om_closest(const rto_scene* s, const rto_bvh* bvh, v3 o, v3 d, float tmin, float tmax,
    float* t, int* prim, float* u, float* v) {
    return om_closest_counted(s, bvh, o, d, tmin, tmax, t, prim, u, v, NULL, NULL);
}

int om_occluded(const rto_scene* s, const rto_bvh* bvh, v3 o, v3 d, float tmin, float tmax) {
    rto_ray r = {{o.x, o.y, o.z}, tmin, {d.x, d.y, d.z}, tmax};
    best_hit b;
    traverse(s, bvh, &r, 1, &b, NULL, NULL);
    return b.found;
}
