// rt_host.h -- host helpers shared by the C ABI implementation.
#pragma once

#include <array>
#include <cstdint>

#include "../../include/rt/rt_hip.h"

namespace rth {

rt_openpbr_material openpbr_defaults();
rt_profile default_profile(int spp, int depth);
void make_lookat_pinhole(rt_camera* c, int w, int h, double vfov_deg, std::array<double, 3> from,
    std::array<double, 3> at, std::array<double, 3> up);
void make_lookat_equi(rt_camera* c, int w, int h, const double radial[6], std::array<double, 3> from,
    std::array<double, 3> at, std::array<double, 3> up);
void make_equi_lut(rt_camera* c);
bool generate_scene(int config, uint64_t seed, rt_generated_scene* g);
bool generate_rig(int config, uint64_t seed, int frame, int frames, rt_camera_rig* out);
void free_generated(rt_generated_scene* g);

}  // namespace rth
