// rt_scenes.cpp -- seeded synthetic benchmark scenes and camera rigs.
//
// The benchmark configurations are new scenes (the reference's builtin scenes
// come from build_scene(id), src/scene/shared_scene_builders.h:56-57, and
// make_rig, utils/render_realtime.cpp:288):
//   0 random_spheres        : ground + 22x22 small spheres + 3 large; one
//                             pinhole 400x225, vfov 20; sky gradient
//   1 cornell_mesh          : 555-unit Cornell room + 130x105 light (15),
//                             icosphere subdivided 6x (81,920 triangles),
//                             one constant OpenPBR material; 4 pinhole
//                             640x480, vfov 40, on a radius-800 ring
//   2 instanced_mesh_restir : 1000 rigid instances of a 1024-triangle mesh
//                             in a 100-unit cube, 256 analytic sphere lights;
//                             2 pinhole32 640x480 + 2 equi62_lut1d 1024x1024
//   3 radiance_cache_realtime : the instanced_mesh_restir scene with 2 pinhole32 + 2
//                             equi62_lut1d cameras at 1280x720, depth 8
//   4 cornell_random        : seeded Cornell rooms with 1-8 spheres, random
//                             materials and light (denoiser training scenes)
//   5 smoke, 6 final_room   : the reference's builtin realtime scenes of the
//                             same names (make_realtime_smoke_scene /
//                             make_final_room_scene,
//                             src/scene/shared_scene_builders.cpp:475-550)
//                             with its four-camera surround rig
//                             (make_default_viewer_rig,
//                             src/realtime/viewer/four_camera_rig.cpp:55-95)
// Draws come from PCG32 (O'Neill 2014; pcg32_srandom_r(seed, 54), XSH-RR
// output); a float draw is (x >> 8) * 2^-24; the draw order is part of the
// scene definition.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <unordered_map>
#include <vector>

#include "../../include/rt/rt_hip.h"
#include "rt_host.h"

namespace rth {

namespace {

class Pcg32 {
   public:
    explicit Pcg32(uint64_t seed) {
        state_ = 0u;
        inc_ = (54u << 1u) | 1u;
        next();
        state_ += seed;
        next();
    }
    uint32_t next() {
        const uint64_t old = state_;
        state_ = old * 6364136223846793005ULL + inc_;
        const uint32_t xs = static_cast<uint32_t>(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = static_cast<uint32_t>(old >> 59u);
        return (xs >> rot) | (xs << ((-rot) & 31u));
    }
    float uniform() { return static_cast<float>(next() >> 8) * (1.0f / 16777216.0f); }

   private:
    uint64_t state_;
    uint64_t inc_;
};

void set3(float* d, float x, float y, float z) {
    d[0] = x;
    d[1] = y;
    d[2] = z;
}

rt_material lambert(float r, float g, float b) {
    rt_material m{};
    m.type = RT_MAT_LAMBERTIAN;
    set3(m.albedo, r, g, b);
    m.ior = 1.0f;
    m.openpbr_index = -1;
    return m;
}

template <typename T>
T* copy_out(const std::vector<T>& v) {
    if (v.empty()) return nullptr;
    T* p = static_cast<T*>(std::malloc(sizeof(T) * v.size()));
    std::memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

void generate_random_spheres(uint64_t seed, rt_generated_scene* g) {
    std::vector<rt_sphere> sp;
    std::vector<rt_material> mt;
    Pcg32 r(seed);
    auto add = [&](float x, float y, float z, float radius, const rt_material& m) {
        rt_sphere s{};
        set3(s.center, x, y, z);
        s.radius = radius;
        s.material = static_cast<int32_t>(mt.size());
        sp.push_back(s);
        mt.push_back(m);
    };
    add(0.0f, -1000.0f, 0.0f, 1000.0f, lambert(0.5f, 0.5f, 0.5f));
    for (int a = -11; a < 11; ++a) {
        for (int b = -11; b < 11; ++b) {
            const float choose = r.uniform();
            const float u = r.uniform();
            const float v = r.uniform();
            rt_material m{};
            m.openpbr_index = -1;
            m.ior = 1.0f;
            if (choose < 0.8f) {
                m.type = RT_MAT_LAMBERTIAN;
                for (int ch = 0; ch < 3; ++ch) {
                    const float x0 = r.uniform();
                    const float x1 = r.uniform();
                    m.albedo[ch] = x0 * x1;
                }
            } else if (choose < 0.95f) {
                m.type = RT_MAT_METAL;
                for (int ch = 0; ch < 3; ++ch) m.albedo[ch] = 0.5f + 0.5f * r.uniform();
                m.fuzz = 0.5f * r.uniform();
            } else {
                m.type = RT_MAT_DIELECTRIC;
                m.ior = 1.5f;
            }
            add(static_cast<float>(a) + 0.9f * u, 0.2f, static_cast<float>(b) + 0.9f * v, 0.2f, m);
        }
    }
    rt_material glass{};
    glass.type = RT_MAT_DIELECTRIC;
    glass.ior = 1.5f;
    glass.openpbr_index = -1;
    add(0.0f, 1.0f, 0.0f, 1.0f, glass);
    add(-4.0f, 1.0f, 0.0f, 1.0f, lambert(0.4f, 0.2f, 0.1f));
    rt_material metal{};
    metal.type = RT_MAT_METAL;
    set3(metal.albedo, 0.7f, 0.6f, 0.5f);
    metal.ior = 1.0f;
    metal.openpbr_index = -1;
    add(4.0f, 1.0f, 0.0f, 1.0f, metal);

    g->scene.sphere_count = static_cast<int32_t>(sp.size());
    g->scene.material_count = static_cast<int32_t>(mt.size());
    g->scene.spheres = copy_out(sp);
    g->scene.materials = copy_out(mt);
    g->scene.background_mode = RT_BG_SKY_GRADIENT;
    set3(g->scene.background, 1.0f, 1.0f, 1.0f);
    set3(g->scene.background_top, 0.5f, 0.7f, 1.0f);
    g->camera_count = 1;
    make_lookat_pinhole(&g->cameras[0], 400, 225, 20.0, {13.0, 2.0, 3.0}, {0.0, 0.0, 0.0},
        {0.0, 1.0, 0.0});
    g->profile = default_profile(32, 10);
}

struct Icosphere {
    std::vector<std::array<double, 3>> verts;
    std::vector<std::array<int32_t, 3>> faces;
};

// unit icosphere with `levels` 4:1 subdivisions, outward winding; new
// vertices are created in face order, edges (a,b), (b,c), (c,a)
Icosphere make_icosphere(int levels) {
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    const double base[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t},
        {0, 1, t}, {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    const int32_t bf[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11},
        {1, 5, 9}, {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2},
        {3, 2, 6}, {3, 6, 8}, {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    Icosphere m;
    for (const auto& v : base) {
        const double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        m.verts.push_back({v[0] / l, v[1] / l, v[2] / l});
    }
    for (const auto& f : bf) {
        int32_t a = f[0], b = f[1], c = f[2];
        double e1[3], e2[3], ce[3];
        for (int k = 0; k < 3; ++k) {
            e1[k] = m.verts[b][k] - m.verts[a][k];
            e2[k] = m.verts[c][k] - m.verts[a][k];
            ce[k] = m.verts[a][k] + m.verts[b][k] + m.verts[c][k];
        }
        const double cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
            e1[0] * e2[1] - e1[1] * e2[0]};
        if (cr[0] * ce[0] + cr[1] * ce[1] + cr[2] * ce[2] < 0.0) std::swap(b, c);
        m.faces.push_back({a, b, c});
    }
    for (int l = 0; l < levels; ++l) {
        std::unordered_map<uint64_t, int32_t> mid;
        mid.reserve(m.faces.size() * 2);
        auto midpoint = [&](int32_t a, int32_t b) {
            const uint64_t lo = static_cast<uint64_t>(a < b ? a : b);
            const uint64_t hi = static_cast<uint64_t>(a < b ? b : a);
            const uint64_t key = (lo << 32) | hi;
            const auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            std::array<double, 3> p;
            for (int k = 0; k < 3; ++k) p[k] = 0.5 * (m.verts[a][k] + m.verts[b][k]);
            const double len = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
            for (int k = 0; k < 3; ++k) p[k] /= len;
            const int32_t id = static_cast<int32_t>(m.verts.size());
            m.verts.push_back(p);
            mid.emplace(key, id);
            return id;
        };
        std::vector<std::array<int32_t, 3>> next;
        next.reserve(m.faces.size() * 4);
        for (const auto& f : m.faces) {
            const int32_t a = f[0], b = f[1], c = f[2];
            const int32_t ab = midpoint(a, b);
            const int32_t bc = midpoint(b, c);
            const int32_t ca = midpoint(c, a);
            next.push_back({a, ab, ca});
            next.push_back({b, bc, ab});
            next.push_back({c, ca, bc});
            next.push_back({ab, bc, ca});
        }
        m.faces.swap(next);
    }
    return m;
}

rt_quad quad(float ox, float oy, float oz, float ux, float uy, float uz, float vx, float vy,
    float vz, int mat) {
    rt_quad q{};
    set3(q.origin, ox, oy, oz);
    set3(q.edge_u, ux, uy, uz);
    set3(q.edge_v, vx, vy, vz);
    q.material = mat;
    return q;
}

void generate_cornell_mesh(uint64_t seed, rt_generated_scene* g) {
    Pcg32 r(seed);
    std::vector<rt_material> mt = {lambert(0.65f, 0.05f, 0.05f), lambert(0.73f, 0.73f, 0.73f),
        lambert(0.12f, 0.45f, 0.15f)};
    rt_material light{};
    light.type = RT_MAT_DIFFUSE_LIGHT;
    set3(light.emission, 15.0f, 15.0f, 15.0f);
    light.ior = 1.0f;
    light.openpbr_index = -1;
    mt.push_back(light);
    rt_openpbr_material op = openpbr_defaults();
    for (int ch = 0; ch < 3; ++ch) op.base_color[ch] = r.uniform();
    op.specular_roughness = 0.05f + 0.85f * r.uniform();
    op.base_metalness = r.uniform() < 0.5f ? 0.0f : 1.0f;
    op.specular_ior = 1.5f;
    op.coat_weight = 0.5f * r.uniform();
    rt_material pbr{};
    pbr.type = RT_MAT_OPENPBR;
    pbr.openpbr_index = 0;
    pbr.ior = 1.0f;
    for (int ch = 0; ch < 3; ++ch) pbr.albedo[ch] = op.base_color[ch];
    mt.push_back(pbr);

    const std::vector<rt_quad> q = {quad(555, 0, 0, 0, 555, 0, 0, 0, 555, 2),
        quad(0, 0, 0, 0, 555, 0, 0, 0, 555, 0), quad(343, 554, 332, -130, 0, 0, 0, 0, -105, 3),
        quad(0, 0, 0, 555, 0, 0, 0, 0, 555, 1), quad(555, 555, 555, -555, 0, 0, 0, 0, -555, 1),
        quad(0, 0, 555, 555, 0, 0, 0, 555, 0, 1)};

    const Icosphere ico = make_icosphere(6);
    const double center[3] = {277.5, 277.5, 277.5};
    const double radius = 150.0;
    std::vector<std::array<float, 3>> pos(ico.verts.size());
    for (size_t i = 0; i < ico.verts.size(); ++i) {
        const float u = r.uniform();
        const double rr = radius * (1.0 + 0.05 * (2.0 * static_cast<double>(u) - 1.0));
        for (int k = 0; k < 3; ++k) pos[i][k] = static_cast<float>(center[k] + ico.verts[i][k] * rr);
    }
    std::vector<rt_triangle> tr(ico.faces.size());
    for (size_t f = 0; f < ico.faces.size(); ++f) {
        std::memcpy(tr[f].p0, pos[ico.faces[f][0]].data(), 12);
        std::memcpy(tr[f].p1, pos[ico.faces[f][1]].data(), 12);
        std::memcpy(tr[f].p2, pos[ico.faces[f][2]].data(), 12);
        tr[f].material = 4;
    }
    g->scene.quad_count = static_cast<int32_t>(q.size());
    g->scene.quads = copy_out(q);
    g->scene.triangle_count = static_cast<int32_t>(tr.size());
    g->scene.triangles = copy_out(tr);
    g->scene.material_count = static_cast<int32_t>(mt.size());
    g->scene.materials = copy_out(mt);
    g->scene.openpbr_count = 1;
    g->scene.openpbr = copy_out(std::vector<rt_openpbr_material>{op});
    g->scene.background_mode = RT_BG_CONSTANT;
    set3(g->scene.background, 0.0f, 0.0f, 0.0f);
    g->camera_count = 4;
    for (int c = 0; c < 4; ++c) {
        const double th = static_cast<double>(c) * 0.5 * 3.14159265358979323846;
        make_lookat_pinhole(&g->cameras[c], 640, 480, 40.0,
            {center[0] + 800.0 * std::sin(th), center[1], center[2] - 800.0 * std::cos(th)},
            {center[0], center[1], center[2]}, {0.0, 1.0, 0.0});
    }
    g->profile = default_profile(16, 8);
}

// cornell_random (id 4): the Cornell room of cornell_mesh with 1-8 displaced icospheres (4
// subdivisions, 5,120 triangles each), one random constant OpenPBR material
// per sphere (drawn as in cornell_mesh), a random ceiling light and four
// cameras in front of the open side.  Draw order:
// light width, depth, x, z, radiance; sphere count; per sphere radius,
// centre x, y, z, base colour, roughness, metalness, coat, then one
// displacement per vertex.
void generate_cornell_random(uint64_t seed, rt_generated_scene* g) {
    Pcg32 r(seed);
    std::vector<rt_material> mt = {lambert(0.65f, 0.05f, 0.05f), lambert(0.73f, 0.73f, 0.73f),
        lambert(0.12f, 0.45f, 0.15f)};
    const float lw = 80.0f + 80.0f * r.uniform();
    const float ld = 70.0f + 70.0f * r.uniform();
    const float lx = 20.0f + (515.0f - lw) * r.uniform();
    const float lz = 20.0f + (515.0f - ld) * r.uniform();
    const float radiance = 8.0f + 12.0f * r.uniform();
    rt_material light{};
    light.type = RT_MAT_DIFFUSE_LIGHT;
    set3(light.emission, radiance, radiance, radiance);
    light.ior = 1.0f;
    light.openpbr_index = -1;
    mt.push_back(light);
    const std::vector<rt_quad> q = {quad(555, 0, 0, 0, 555, 0, 0, 0, 555, 2),
        quad(0, 0, 0, 0, 555, 0, 0, 0, 555, 0), quad(lx + lw, 554, lz + ld, -lw, 0, 0, 0, 0, -ld, 3),
        quad(0, 0, 0, 555, 0, 0, 0, 0, 555, 1), quad(555, 555, 555, -555, 0, 0, 0, 0, -555, 1),
        quad(0, 0, 555, 555, 0, 0, 0, 555, 0, 1)};
    const int spheres = 1 + std::min(7, static_cast<int>(8.0f * r.uniform()));
    const Icosphere ico = make_icosphere(4);
    std::vector<rt_openpbr_material> ops;
    std::vector<rt_triangle> tr;
    for (int m = 0; m < spheres; ++m) {
        const double radius = 40.0 + 70.0 * static_cast<double>(r.uniform());
        double center[3];
        for (int k = 0; k < 3; ++k) center[k] = radius + (555.0 - 2.0 * radius) * static_cast<double>(r.uniform());
        rt_openpbr_material op = openpbr_defaults();
        for (int ch = 0; ch < 3; ++ch) op.base_color[ch] = r.uniform();
        op.specular_roughness = 0.05f + 0.85f * r.uniform();
        op.base_metalness = r.uniform() < 0.5f ? 0.0f : 1.0f;
        op.specular_ior = 1.5f;
        op.coat_weight = 0.5f * r.uniform();
        rt_material pbr{};
        pbr.type = RT_MAT_OPENPBR;
        pbr.openpbr_index = m;
        pbr.ior = 1.0f;
        for (int ch = 0; ch < 3; ++ch) pbr.albedo[ch] = op.base_color[ch];
        mt.push_back(pbr);
        ops.push_back(op);
        std::vector<std::array<float, 3>> pos(ico.verts.size());
        for (size_t i = 0; i < ico.verts.size(); ++i) {
            const float u = r.uniform();
            const double rr = radius * (1.0 + 0.05 * (2.0 * static_cast<double>(u) - 1.0));
            for (int k = 0; k < 3; ++k) pos[i][k] = static_cast<float>(center[k] + ico.verts[i][k] * rr);
        }
        for (size_t f = 0; f < ico.faces.size(); ++f) {
            rt_triangle t{};
            std::memcpy(t.p0, pos[ico.faces[f][0]].data(), 12);
            std::memcpy(t.p1, pos[ico.faces[f][1]].data(), 12);
            std::memcpy(t.p2, pos[ico.faces[f][2]].data(), 12);
            t.material = 4 + m;
            tr.push_back(t);
        }
    }
    g->scene.quad_count = static_cast<int32_t>(q.size());
    g->scene.quads = copy_out(q);
    g->scene.triangle_count = static_cast<int32_t>(tr.size());
    g->scene.triangles = copy_out(tr);
    g->scene.material_count = static_cast<int32_t>(mt.size());
    g->scene.materials = copy_out(mt);
    g->scene.openpbr_count = static_cast<int32_t>(ops.size());
    g->scene.openpbr = copy_out(ops);
    g->scene.background_mode = RT_BG_CONSTANT;
    set3(g->scene.background, 0.0f, 0.0f, 0.0f);
    // four cameras on an arc in front of the open side (azimuth -24, -8, 8,
    // 24 degrees at radius 800, heights +100, -60, +60, -100), all aimed at
    // the room centre: every image sees the lit interior
    g->camera_count = 4;
    const double c0[3] = {277.5, 277.5, 277.5};
    const double azimuth[4] = {-24.0, -8.0, 8.0, 24.0}, height[4] = {100.0, -60.0, 60.0, -100.0};
    for (int c = 0; c < 4; ++c) {
        const double th = azimuth[c] * (3.14159265358979323846 / 180.0);
        make_lookat_pinhole(&g->cameras[c], 640, 480, 40.0,
            {c0[0] + 800.0 * std::sin(th), c0[1] + height[c], c0[2] - 800.0 * std::cos(th)},
            {c0[0], c0[1], c0[2]}, {0.0, 1.0, 0.0});
    }
    g->profile = default_profile(16, 8);
}

double gauss(Pcg32& r, double sigma) {
    const double u1 = 1.0 - static_cast<double>(r.uniform());
    const double u2 = static_cast<double>(r.uniform());
    return sigma * std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * 3.14159265358979323846 * u2);
}

// four cameras on a radius-70 orbit around the cube centre, 90 degrees apart:
// two pinhole32 640x480 (vfov 60) and two equi62_lut1d 1024x1024 with
// distortion k_i ~ N(0, 0.01) redrawn until theta_d is monotone on [0, 1.6]
void instanced_rig(uint64_t seed, int frame, int frames, rt_camera* cams, int wp = 640,
    int hp = 480, int we = 1024, int he = 1024) {
    Pcg32 r(seed ^ 0x5bd1e995ULL);
    double radial[2][6];
    for (int k = 0; k < 2; ++k) {
        for (int tries = 0; tries < 10000; ++tries) {
            for (int i = 0; i < 6; ++i) radial[k][i] = gauss(r, 0.01);
            bool ok = true;
            for (int s = 0; s <= 64 && ok; ++s) {
                const double th = 1.6 * static_cast<double>(s) / 64.0, t2 = th * th;
                double der = 1.0, tp = t2;
                for (int i = 0; i < 6; ++i) {
                    der += static_cast<double>(2 * i + 3) * radial[k][i] * tp;
                    tp *= t2;
                }
                ok = der >= 0.2;
            }
            if (ok) break;
        }
    }
    const std::array<double, 3> center = {50.0, 50.0, 50.0};
    const std::array<double, 3> up = {0.0, 1.0, 0.0};
    for (int k = 0; k < 4; ++k) {
        const double th = 2.0 * 3.14159265358979323846 * static_cast<double>(frame)
                              / static_cast<double>(frames > 0 ? frames : 1)
                          + static_cast<double>(k) * 0.5 * 3.14159265358979323846;
        const std::array<double, 3> from = {center[0] + 70.0 * std::cos(th), center[1] + 15.0,
            center[2] + 70.0 * std::sin(th)};
        if (k < 2) {
            make_lookat_pinhole(&cams[k], wp, hp, 60.0, from, center, up);
        } else {
            make_lookat_equi(&cams[k], we, he, radial[k - 2], from, center, up);
        }
    }
}

void generate_instanced_mesh(uint64_t seed, rt_generated_scene* g) {
    Pcg32 r(seed);
    constexpr int NU = 32, NV = 16, NINST = 1000, NMAT = 16, NLIGHT = 256;
    constexpr double kPiD = 3.14159265358979323846;
    // torus, 32 x 16 segments = 1024 triangles, tube radius jittered per vertex
    std::vector<std::array<double, 3>> verts(NU * NV);
    for (int i = 0; i < NU; ++i) {
        for (int j = 0; j < NV; ++j) {
            const double u = 2.0 * kPiD * static_cast<double>(i) / NU;
            const double v = 2.0 * kPiD * static_cast<double>(j) / NV;
            const double tube = 0.35 * (1.0 + 0.1 * (2.0 * static_cast<double>(r.uniform()) - 1.0));
            verts[i * NV + j] = {(1.0 + tube * std::cos(v)) * std::cos(u), tube * std::sin(v),
                (1.0 + tube * std::cos(v)) * std::sin(u)};
        }
    }
    std::vector<std::array<int, 3>> faces;
    for (int i = 0; i < NU; ++i) {
        for (int j = 0; j < NV; ++j) {
            const int a = i * NV + j, b = ((i + 1) % NU) * NV + j;
            const int c = ((i + 1) % NU) * NV + (j + 1) % NV, d = i * NV + (j + 1) % NV;
            faces.push_back({a, b, c});
            faces.push_back({a, c, d});
        }
    }
    std::vector<rt_material> mt(NMAT);
    for (auto& m : mt) {
        m = rt_material{};
        m.type = RT_MAT_LAMBERTIAN;
        m.ior = 1.0f;
        m.openpbr_index = -1;
        for (int ch = 0; ch < 3; ++ch) m.albedo[ch] = 0.2f + 0.7f * r.uniform();
    }
    std::vector<rt_triangle> tr((size_t)NINST * faces.size());
    for (int k = 0; k < NINST; ++k) {
        // uniform rotation from a random unit quaternion (Shoemake 1992)
        const double u1 = r.uniform(), u2 = r.uniform(), u3 = r.uniform();
        const double s1 = std::sqrt(1.0 - u1), s2 = std::sqrt(u1);
        const double qx = s1 * std::sin(2.0 * kPiD * u2);
        const double qy = s1 * std::cos(2.0 * kPiD * u2);
        const double qz = s2 * std::sin(2.0 * kPiD * u3);
        const double qw = s2 * std::cos(2.0 * kPiD * u3);
        const double R[3][3] = {
            {1.0 - 2.0 * (qy * qy + qz * qz), 2.0 * (qx * qy - qz * qw), 2.0 * (qx * qz + qy * qw)},
            {2.0 * (qx * qy + qz * qw), 1.0 - 2.0 * (qx * qx + qz * qz), 2.0 * (qy * qz - qx * qw)},
            {2.0 * (qx * qz - qy * qw), 2.0 * (qy * qz + qx * qw), 1.0 - 2.0 * (qx * qx + qy * qy)}};
        const double t[3] = {100.0 * static_cast<double>(r.uniform()),
            100.0 * static_cast<double>(r.uniform()), 100.0 * static_cast<double>(r.uniform())};
        const int mat = static_cast<int>(r.uniform() * static_cast<float>(NMAT)) % NMAT;
        const double scale = 2.0;
        for (size_t f = 0; f < faces.size(); ++f) {
            rt_triangle& o = tr[(size_t)k * faces.size() + f];
            float* dst[3] = {o.p0, o.p1, o.p2};
            for (int vi = 0; vi < 3; ++vi) {
                const auto& p = verts[faces[f][vi]];
                for (int a = 0; a < 3; ++a)
                    dst[vi][a] = static_cast<float>(
                        scale * (R[a][0] * p[0] + R[a][1] * p[1] + R[a][2] * p[2]) + t[a]);
            }
            o.material = mat;
        }
    }
    std::vector<rt_analytic_light> al(NLIGHT);
    for (auto& L : al) {
        L = rt_analytic_light{};
        L.type = RT_ALIGHT_SPHERE;
        for (int a = 0; a < 3; ++a) L.position[a] = 100.0f * r.uniform();
        L.radius = 0.1f + 0.4f * r.uniform();
        const float rad = 1.0f + 19.0f * r.uniform();
        set3(L.radiance, rad, rad, rad);
        set3(L.basis_x, 1.0f, 0.0f, 0.0f);
        set3(L.basis_y, 0.0f, 1.0f, 0.0f);
        set3(L.basis_z, 0.0f, 0.0f, 1.0f);
        L.width = L.height = L.length = 1.0f;
        L.cos_theta_max = 1.0f;
    }
    // world areas and the power x area selection distribution (identity
    // bases are similarity transforms, so the compile cannot refuse them)
    (void)rt_analytic_lights_compile(al.data(), NLIGHT);
    g->scene.triangle_count = static_cast<int32_t>(tr.size());
    g->scene.triangles = copy_out(tr);
    g->scene.material_count = NMAT;
    g->scene.materials = copy_out(mt);
    g->scene.analytic_light_count = NLIGHT;
    g->scene.analytic_lights = copy_out(al);
    g->scene.background_mode = RT_BG_CONSTANT;
    set3(g->scene.background, 0.0f, 0.0f, 0.0f);
    g->camera_count = 4;
    instanced_rig(seed, 0, 60, g->cameras);
    g->profile = default_profile(1, 2);
    g->profile.enable_restir_di = 1;
    g->profile.restir_initial_candidates = 32;
    g->profile.restir_spatial_neighbors = 5;
    g->profile.restir_spatial_radius = 30;
}

}  // namespace

// equi62 theta distortion and its Newton inverse (camera_models.cpp:64-100)
static double equi_theta_dist(const double* radial, double theta) {
    const double t2 = theta * theta;
    double scale = 1.0, tp = t2;
    for (int i = 0; i < 6; ++i) {
        scale += radial[i] * tp;
        tp *= t2;
    }
    return theta * scale;
}

static double equi_theta_deriv(const double* radial, double theta) {
    const double t2 = theta * theta;
    double der = 1.0, tp = t2;
    for (int i = 0; i < 6; ++i) {
        der += static_cast<double>(2 * i + 3) * radial[i] * tp;
        tp *= t2;
    }
    return der;
}

static double equi_invert(const double* radial, double thetad) {
    double theta = std::cbrt(thetad > 0.0 ? thetad : 0.0);
    for (int it = 0; it < 30; ++it) {
        const double err = thetad - equi_theta_dist(radial, theta);
        if (std::fabs(err) < 1e-12) break;
        const double der = equi_theta_deriv(radial, theta);
        if (std::fabs(der) < 1e-12) break;
        const double delta = err / der;
        theta += delta;
        if (std::fabs(delta) < 1e-12 || std::fabs(delta) > 3.1416) break;
    }
    return theta;
}

// make_equi62_lut1d_params (camera_models.cpp:228-248)
void make_equi_lut(rt_camera* c) {
    const double x = (static_cast<double>(c->width) + 10.0 - c->cx) / c->fx;
    const double y = (static_cast<double>(c->height) + 10.0 - c->cy) / c->fy;
    c->lut_step = std::sqrt(x * x + y * y) / 1024.0;
    for (int i = 0; i < 1024; ++i) c->lut[i] = equi_invert(c->radial, c->lut_step * static_cast<double>(i));
}

// equi62_lut1d look-at camera with the default 120 degree horizontal field of
// view (derive_default_camera_intrinsics, camera_models.cpp:196-226)
void make_lookat_equi(rt_camera* c, int w, int h, const double radial[6],
    std::array<double, 3> from, std::array<double, 3> at, std::array<double, 3> up) {
    make_lookat_pinhole(c, w, h, 60.0, from, at, up);
    c->model = RT_CAM_EQUI62_LUT1D;
    const double hfov = 120.0 * 3.14159265358979323846 / 180.0;
    const double f = 0.5 * static_cast<double>(w) / (0.5 * hfov);
    c->fx = f;
    c->fy = f;
    c->cx = 0.5 * static_cast<double>(w);
    c->cy = 0.5 * static_cast<double>(h);
    for (int i = 0; i < 6; ++i) c->radial[i] = radial[i];
    c->tangential[0] = 0.0;
    c->tangential[1] = 0.0;
    make_equi_lut(c);
}

// The reference's surround rig for a realtime view preset: the body pose's
// camera at yaw offsets 0, 90, -90, 180 degrees (kDefaultSurroundYawOffsetsDeg,
// src/realtime/default_viewer_conventions.h:7); per camera the axes are
// right, -up, forward of the yawed body (forward_direction / right_direction,
// src/realtime/viewer/body_pose.cpp:36-85, world z up), at the body position.
void surround_rig(const double pos[3], double yaw_deg, double pitch_deg, double fx, double fy, double cx,
    double cy, int w, int h, rt_camera* cams) {
    const double offsets[4] = {0.0, 90.0, -90.0, 180.0};
    const double kPi = 3.14159265358979323846;
    for (int i = 0; i < 4; ++i) {
        rt_camera* c = &cams[i];
        std::memset(c, 0, sizeof(*c));
        c->enabled = 1;
        c->width = w;
        c->height = h;
        c->model = RT_CAM_PINHOLE32;
        const double yaw = (yaw_deg + offsets[i]) * kPi / 180.0;
        const double pitch = -pitch_deg * kPi / 180.0;
        // body rotation Rx(yaw) * Ry(-pitch) applied to (0, 0, -1)
        const double b[3] = {-std::sin(pitch), std::sin(yaw) * std::cos(pitch),
            -std::cos(yaw) * std::cos(pitch)};
        // body basis (world z up): columns up (0,0,1), -right (-1,0,0), -forward (0,-1,0)
        double f[3] = {-b[1], -b[2], b[0]};
        const double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
        for (double& v : f) v /= fl;
        // right = forward x up
        double r[3] = {f[1], -f[0], 0.0};
        const double rl = std::sqrt(r[0] * r[0] + r[1] * r[1]);
        if (rl * rl < 1e-12) {
            // looking straight up or down: the rotated body -y axis, (cos yaw, sin yaw, 0)
            r[0] = std::cos(yaw);
            r[1] = std::sin(yaw);
            r[2] = 0.0;
        } else {
            for (double& v : r) v /= rl;
        }
        double u[3] = {r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0]};
        const double ul = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        for (double& v : u) v /= ul;
        for (int k = 0; k < 3; ++k) {
            c->origin[k] = pos[k];
            c->basis_x[k] = r[k];
            c->basis_y[k] = -u[k];
            c->basis_z[k] = f[k];
        }
        c->fx = fx;
        c->fy = fy;
        c->cx = cx;
        c->cy = cy;
    }
}

// legacy_renderer_to_world (src/realtime/frame_convention.h:35-37): (x, y, z) -> (x, -z, y)
rt_quad legacy_quad(double ox, double oy, double oz, double ux, double uy, double uz, double vx,
    double vy, double vz, int mat) {
    return quad((float)ox, (float)-oz, (float)oy, (float)ux, (float)-uz, (float)uy, (float)vx,
        (float)-vz, (float)vy, mat);
}

rt_sphere legacy_sphere(double x, double y, double z, double radius, int mat) {
    rt_sphere sp{};
    set3(sp.center, (float)x, (float)-z, (float)y);
    sp.radius = (float)radius;
    sp.material = mat;
    return sp;
}

rt_material emissive(float r, float g, float b) {
    rt_material m{};
    m.type = RT_MAT_DIFFUSE_LIGHT;
    set3(m.emission, r, g, b);
    m.ior = 1.0f;
    m.openpbr_index = -1;
    return m;
}

void finish_builtin(rt_generated_scene* g, const std::vector<rt_sphere>& sp,
    const std::vector<rt_quad>& q, const std::vector<rt_material>& mt) {
    g->scene.sphere_count = static_cast<int32_t>(sp.size());
    g->scene.spheres = sp.empty() ? nullptr : copy_out(sp);
    g->scene.quad_count = static_cast<int32_t>(q.size());
    g->scene.quads = copy_out(q);
    g->scene.material_count = static_cast<int32_t>(mt.size());
    g->scene.materials = copy_out(mt);
    g->scene.background_mode = RT_BG_CONSTANT;
    set3(g->scene.background, 0.0f, 0.0f, 0.0f);
    g->camera_count = 4;
    g->profile = default_profile(16, 8);
}

// make_realtime_smoke_scene (shared_scene_builders.cpp:475-487) and its view
// preset: body at (0, -2.5, 0.25), yaw 0, pinhole 640x480, fx = fy = 360
void generate_smoke(rt_generated_scene* g) {
    const std::vector<rt_material> mt = {lambert(0.75f, 0.25f, 0.2f), emissive(10.0f, 10.0f, 10.0f)};
    const std::vector<rt_sphere> sp = {legacy_sphere(0.0, 0.0, -1.0, 0.5, 0)};
    const std::vector<rt_quad> q = {legacy_quad(-0.75, 1.25, -1.5, 1.5, 0.0, 0.0, 0.0, 0.0, 1.5, 1)};
    finish_builtin(g, sp, q, mt);
    const double pos[3] = {0.0, -2.5, 0.25};
    surround_rig(pos, 0.0, 0.0, 360.0, 360.0, 320.0, 240.0, 640, 480, g->cameras);
}

// make_final_room_scene (shared_scene_builders.cpp:489-550) and its view
// preset: the default spawn pose (0, -0.8, 0.35), yaw 0, pinhole 640x480,
// fx = 0.75 * 640, fy = 0.75 * 480
void generate_final_room(rt_generated_scene* g) {
    const std::vector<rt_material> mt = {lambert(0.73f, 0.73f, 0.73f), lambert(0.30f, 0.70f, 0.35f),
        lambert(0.72f, 0.25f, 0.22f), lambert(0.25f, 0.35f, 0.75f), emissive(10.0f, 10.0f, 10.0f)};
    const int white = 0, green = 1, red = 2, blue = 3, light = 4;
    const std::vector<rt_quad> q = {
        legacy_quad(-4.0, -1.0, -4.0, 0.0, 0.0, 8.0, 8.0, 0.0, 0.0, white),
        legacy_quad(-4.0, 3.5, -4.0, 8.0, 0.0, 0.0, 0.0, 0.0, 8.0, white),
        legacy_quad(-4.0, -1.0, -4.0, 0.0, 4.5, 0.0, 0.0, 0.0, 8.0, green),
        legacy_quad(4.0, -1.0, -4.0, 0.0, 0.0, 8.0, 0.0, 4.5, 0.0, red),
        legacy_quad(-4.0, -1.0, -4.0, 8.0, 0.0, 0.0, 0.0, 4.5, 0.0, white),
        legacy_quad(-4.0, -1.0, 4.0, 0.0, 4.5, 0.0, 8.0, 0.0, 0.0, blue),
        legacy_quad(-1.0, 3.15, -1.0, 2.0, 0.0, 0.0, 0.0, 0.0, 2.0, light),
        legacy_quad(-3.2, -0.25, -3.0, 0.0, 0.0, 1.8, 1.8, 0.0, 0.0, white),
        legacy_quad(1.2, 0.15, 1.0, 0.0, 0.0, 1.6, 1.6, 0.0, 0.0, white)};
    const std::vector<rt_sphere> sp = {legacy_sphere(0.0, 0.25, -1.2, 0.55, white),
        legacy_sphere(-1.6, 0.35, 1.7, 0.55, white), legacy_sphere(-3.1, 1.0, 0.8, 0.55, white),
        legacy_sphere(3.0, 1.35, -0.9, 0.65, white), legacy_sphere(1.1, 1.1, -3.0, 0.60, white),
        legacy_sphere(-0.8, 2.55, 2.2, 0.45, white), legacy_sphere(0.9, 0.55, -0.1, 0.45, white),
        legacy_sphere(-0.35, 0.4, -1.15, 0.35, white)};
    finish_builtin(g, sp, q, mt);
    const double pos[3] = {0.0, -0.8, 0.35};
    surround_rig(pos, 0.0, 0.0, 480.0, 360.0, 320.0, 240.0, 640, 480, g->cameras);
}

bool generate_rig(int config, uint64_t seed, int frame, int frames, rt_camera_rig* out) {
    std::memset(out, 0, sizeof(*out));
    if (config == 2) {
        instanced_rig(seed, frame, frames, out->cameras);
        out->active_count = 4;
        return true;
    }
    if (config == 3) {
        instanced_rig(seed, frame, frames, out->cameras, 1280, 720, 1280, 720);
        out->active_count = 4;
        return true;
    }
    rt_generated_scene g;
    if (!generate_scene(config, seed, &g)) return false;
    out->active_count = g.camera_count;
    for (int c = 0; c < g.camera_count; ++c) out->cameras[c] = g.cameras[c];
    free_generated(&g);
    return true;
}

rt_openpbr_material openpbr_defaults() {
    rt_openpbr_material m{};
    m.base_weight = 1.0f;
    set3(m.base_color, 0.8f, 0.8f, 0.8f);
    m.specular_weight = 1.0f;
    set3(m.specular_color, 1.0f, 1.0f, 1.0f);
    m.specular_roughness = 0.3f;
    m.specular_ior = 1.5f;
    set3(m.transmission_color, 1.0f, 1.0f, 1.0f);
    m.transmission_dispersion_abbe_number = 20.0f;
    set3(m.subsurface_color, 0.8f, 0.8f, 0.8f);
    m.subsurface_radius = 1.0f;
    set3(m.subsurface_radius_scale, 1.0f, 0.5f, 0.25f);
    set3(m.fuzz_color, 1.0f, 1.0f, 1.0f);
    m.fuzz_roughness = 0.5f;
    set3(m.coat_color, 1.0f, 1.0f, 1.0f);
    m.coat_ior = 1.6f;
    m.coat_darkening = 1.0f;
    m.thin_film_thickness = 0.5f;
    m.thin_film_ior = 1.4f;
    set3(m.emission_color, 1.0f, 1.0f, 1.0f);
    m.geometry_opacity = 1.0f;
    return m;
}

rt_profile default_profile(int spp, int depth) {
    rt_profile p{};
    p.samples_per_pixel = spp;
    p.max_bounces = depth;
    p.enable_denoise = 0;
    p.rr_start_bounce = depth;
    p.accumulation_reset_rotation_deg = 2.0;
    p.accumulation_reset_translation = 0.05;
    p.restir_initial_candidates = 4;
    p.restir_temporal_reuse = 1;
    p.restir_max_history_age = 20;
    p.restir_max_temporal_candidates = 64;
    p.restir_spatial_neighbors = 1;
    p.restir_max_spatial_candidates = 4;
    p.restir_bias_correction = 1;
    p.restir_min_analytic_lights = 16;
    p.pixel_jitter = 1;
    p.restir_spatial_radius = 30;
    return p;
}

// OpenCV camera axes: x right, y down, z forward
void make_lookat_pinhole(rt_camera* c, int w, int h, double vfov_deg, std::array<double, 3> from,
    std::array<double, 3> at, std::array<double, 3> up) {
    std::memset(c, 0, sizeof(*c));
    c->enabled = 1;
    c->width = w;
    c->height = h;
    c->model = RT_CAM_PINHOLE32;
    double z[3] = {at[0] - from[0], at[1] - from[1], at[2] - from[2]};
    const double zl = std::sqrt(z[0] * z[0] + z[1] * z[1] + z[2] * z[2]);
    for (double& v : z) v /= zl;
    double x[3] = {z[1] * up[2] - z[2] * up[1], z[2] * up[0] - z[0] * up[2],
        z[0] * up[1] - z[1] * up[0]};
    const double xl = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    for (double& v : x) v /= xl;
    const double y[3] = {z[1] * x[2] - z[2] * x[1], z[2] * x[0] - z[0] * x[2],
        z[0] * x[1] - z[1] * x[0]};
    for (int i = 0; i < 3; ++i) {
        c->origin[i] = from[i];
        c->basis_x[i] = x[i];
        c->basis_y[i] = y[i];
        c->basis_z[i] = z[i];
    }
    const double vfov = vfov_deg * 3.14159265358979323846 / 180.0;
    const double f = 0.5 * static_cast<double>(h) / std::tan(0.5 * vfov);
    c->fx = f;
    c->fy = f;
    c->cx = 0.5 * static_cast<double>(w);
    c->cy = 0.5 * static_cast<double>(h);
}

bool generate_scene(int config, uint64_t seed, rt_generated_scene* g) {
    std::memset(g, 0, sizeof(*g));
    switch (config) {
        case 0: generate_random_spheres(seed, g); return true;
        case 1: generate_cornell_mesh(seed, g); return true;
        case 2: generate_instanced_mesh(seed, g); return true;
        case 3:
            // the instanced_mesh_restir scene, four 1280x720 cameras, paths of depth 8 with
            // Russian roulette from bounce 3 (the radiance-cache workload)
            generate_instanced_mesh(seed, g);
            instanced_rig(seed, 0, 60, g->cameras, 1280, 720, 1280, 720);
            g->profile.max_bounces = 8;
            g->profile.rr_start_bounce = 3;
            return true;
        case 4: generate_cornell_random(seed, g); return true;
        case 5: generate_smoke(g); return true;
        case 6: generate_final_room(g); return true;
        default: return false;
    }
}

void free_generated(rt_generated_scene* g) {
    std::free(const_cast<rt_sphere*>(g->scene.spheres));
    std::free(const_cast<rt_quad*>(g->scene.quads));
    std::free(const_cast<rt_triangle*>(g->scene.triangles));
    std::free(const_cast<rt_material*>(g->scene.materials));
    std::free(const_cast<rt_openpbr_material*>(g->scene.openpbr));
    std::free(const_cast<rt_analytic_light*>(g->scene.analytic_lights));
    std::memset(g, 0, sizeof(*g));
}

}  // namespace rth
