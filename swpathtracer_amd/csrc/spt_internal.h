// spt_internal.h - shared declarations of the core's translation units.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "spt.h"
#include "spt_math.h"

namespace spt {

// Accessors of the host-built tree (bvh_build.cpp).
const std::vector<spt_node64>& bvh_nodes(const spt_bvh* b);
const std::vector<spt_tri48>& bvh_tris(const spt_bvh* b);
const std::vector<uint32_t>& bvh_leaf_prim(const spt_bvh* b);
const float* bvh_root_box(const spt_bvh* b);
// SPT_PRIM_* of every scene primitive; empty when all are triangles.
const std::vector<uint8_t>& bvh_kinds(const spt_bvh* b);

// Camera state precomputed on the host exactly as the reference's camera
// constructors do (cameras/camera.cpp:3-11, pinhole_camera.cpp:3-12,
// orthographic_camera.cpp:5-19); rays are formed per sample on the device.
struct DevCamera {
    uint32_t type;
    uint32_t width, height;
    uint32_t pad;
    V3 position;
    Frame frame;  // local-to-world columns
    V3 ulc;       // local upper-left corner
    float pdx, pdy;
    V3 ortho_dir;
};

DevCamera make_camera(const spt_camera& c);

}  // namespace spt
