// Host-side conversion of a reference-layout scene (rm_scene_desc) into the
// device layouts of rm_layout.h. Plain C++: no HIP calls, so it can be checked
// on a machine without a GPU.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "rm_hip.h"

namespace rmh {

struct F4 {
    float x, y, z, w;
};

struct PackedScene {
    std::vector<F4> nodes;        // kNodeWords per 4-wide node, preorder (parents before children)
    std::vector<int32_t> depth;   // tree depth of each 4-wide node
    std::vector<F4> geo;          // kTriGeoWords per triangle, leaf order
    std::vector<F4> nrm;          // kTriNrmWords per triangle, scene order
    std::vector<F4> nn;           // 1 per triangle
    std::vector<int32_t> matid;   // 1 per triangle
    std::vector<F4> mats;         // kMatWords per material
    std::vector<F4> uv;           // 2 per triangle when any triangle is textured, else empty
    std::vector<F4> lights;       // kLightWords per light, CDF order
    std::vector<float> cdf;
    std::vector<float> texels;
    float light_total = 0.0f;
    int32_t opaque = 1;
    int32_t stack_need = 0;       // deepest possible traversal stack
};

// Validates the arrays and fills `out`. Returns false with a message in `err`
// when the scene is malformed (indices out of range, inconsistent tree).
bool pack_scene(const rm_scene_desc& s, PackedScene& out, std::string& err);

// Same triangle count and identical tree topology (escape links, leaf ranges
// and triangle order), i.e. the resident 4-wide tree can be refit in place.
bool same_topology(const rm_scene_desc& a_desc, const std::vector<int32_t>& prev_bvh_tri,
                   const std::vector<int32_t>& prev_topo);
void topology_of(const rm_scene_desc& s, std::vector<int32_t>& bvh_tri, std::vector<int32_t>& topo);

// 64-bit content hashes (8 bytes per step).
uint64_t hash_bytes(const void* data, size_t n, uint64_t h);
// Temporal-history key: triangle records, triangle and light counts and the
// light total, the fields metal.mm:241-248 covers.
uint64_t history_key(const rm_scene_desc& s);
// Everything that is uploaded.
uint64_t content_key(const rm_scene_desc& s);

}  // namespace rmh
