// TEST INFRASTRUCTURE ONLY. The CPU oracle checks the HIP backend; nothing in
// the shipped library includes, links or calls it.
//
// Byte layouts of the records the reference's Metal backend uploads
// (src/engine/metal_types.h:3-57, sizes asserted at src/engine/metal.mm:13-17).
#pragma once
#include <cstddef>
#include <cstdint>

extern "C" {

typedef struct { float x, y, z; } orc_vec3;

typedef struct {  // metal_types.h:5-11, 176 bytes
    orc_vec3 v0, v1, v2, n, nn, n0, n1, n2;
    float ar, ag, ab, er, eg, eb, ior, rough, metallic, alpha;
    int32_t shader;
    float tu0, tv0, tu1, tv1, tu2, tv2;
    int32_t texture_offset, texture_width, texture_height;
} orc_triangle;

typedef struct {  // metal_types.h:12-15, 52 bytes
    float x, y, z, focal;
    float m0, m1, m2, m3, m4, m5, m6, m7, m8;
} orc_camera;

typedef struct {  // metal_types.h:17-20, 36 bytes; right = preorder escape index
    orc_vec3 mn, mx;
    int32_t right, offset, count;
} orc_node;

typedef struct {  // metal_types.h:21-30, 184 bytes
    orc_camera camera;
    int32_t width, height, spp, bounces, nodes, lights, adapt_min, adapt_step, nomis, opaque;
    float light_total, resolution, clamp_value, adapt_rel, adapt_abs;
    int32_t restir_candidates, restir_spatial, restir_mcap;
    float restir_radius;
    orc_camera prev_camera;
    int32_t has_history;
} orc_params;

typedef struct {  // metal_types.h:32-42, 76 bytes
    orc_vec3 pos, normal, albedo, wo, emission;
    float rough, metallic, ior;
    int32_t kind;  // 0 miss, 1 opaque, 3 dielectric
} orc_gbuffer;

typedef struct {  // metal_types.h:46-56, 60 bytes
    orc_vec3 light_pos, light_normal, emission;
    float alpha, p_hat, sum_ws, sum_wu;
    uint32_t m, valid;
} orc_reservoir;

// A scene as uploaded to the device (metal.mm:109-145).
typedef struct {
    const orc_triangle* tris; size_t ntri;
    const orc_node* nodes; size_t nnode;
    const int32_t* indices;          // bvh_tri
    const int32_t* lights; const float* cdf; size_t nlight;
    float light_total;
    const float* texels; size_t ntexel;  // RGB triples
} orc_scene;

// Render settings after the host-side defaults and environment handling of
// metal.mm:146-170 (what the shim passes to the backend).
typedef struct {
    int32_t width, height, spp, bounces;
    float resolution, clamp_value;
    int32_t adapt_min, adapt_step;
    float adapt_rel, adapt_abs;
    int32_t nomis;
    int32_t restir_candidates, restir_spatial, restir_mcap;
    float restir_radius;
    int32_t denoise, denoise_iterations, temporal;
    orc_camera camera;
} orc_frame;

}  // extern "C"
