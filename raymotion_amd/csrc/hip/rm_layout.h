// Device-resident data layouts of the MI355X backend.
//
// The reference keeps one 176-byte record per triangle and walks a binary
// preorder tree (metal_types.h). Here the data are split by who reads them:
//   traversal  : 4-wide BVH nodes (128 B = one L2 line) and triangle vertices
//                in leaf order (48 B, the only per-triangle data a traversal
//                step touches);
//   shading    : vertex + geometric normals per triangle (48 B), a material
//                index per triangle into a deduplicated material table, UVs
//                only for textured triangles, the unit geometric normal only
//                for emission MIS;
//   lights     : one 64-B record per emissive triangle in CDF order;
//   per pixel  : G-buffer and reservoirs as structure-of-arrays of float4 so
//                every pass reads and writes fully coalesced 16-B lanes.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "rm_hip.h"

namespace rmh {

// 4-wide node: 8 x float4. Words 0..5 hold the child boxes as
// lo.x[4], hi.x[4], lo.y[4], hi.y[4], lo.z[4], hi.z[4]; word 6 the child
// references and word 7 the child kinds (int bits):
//   kind == -1 empty slot, kind == 0 internal child (ref = node index),
//   kind  >  0 leaf with `kind` triangles starting at leaf-order index ref.
constexpr int kNodeWords = 8;
constexpr int kTriGeoWords = 3;   // [v0.xyz v1.x] [v1.yz v2.xy] [v2.z id - -]
constexpr int kTriNrmWords = 3;   // [n0.xyz n.x] [n1.xyz n.y] [n2.xyz n.z]
constexpr int kMatWords = 4;      // [ar ag ab alpha] [er eg eb ior] [rough metallic shader texoff] [texw texh - -]
constexpr int kLightWords = 4;    // [v0.xyz v1.x] [v1.yz v2.xy] [v2.z nn.xyz] [er eg eb alpha]
constexpr int kBlock = 256;       // threads per block of every pixel pass

// Per-pass counters: counters[3 * pass + {0 rays, 1 node visits, 2 triangle
// tests}], then a stack-overflow flag. Lanes add once per kernel; the
// compiler folds uniform-address atomics into one per wave.
enum { kCntMega = 0, kCntGenerate, kCntSpatial, kCntShade, kCntProbe, kNumPasses };
constexpr int kCntOverflow = 3 * kNumPasses;
constexpr int kNumCounters = kCntOverflow + 1;

struct DevScene {
    const float4* nodes;
    const float4* geo;
    const float4* nrm;
    const float4* nn;
    const int32_t* matid;
    const float4* mats;
    const float4* uv;
    const float4* lights;
    const float* cdf;
    const float* texels;
    int32_t nnodes, nlights, opaque;
    float light_total;
};

struct FrameParams {
    rm_camera cam, prev;
    int32_t width, height, spp, bounces;
    int32_t adapt_min, adapt_step, nomis, has_history;
    float resolution, clamp_value, adapt_rel, adapt_abs;
    int32_t restir_candidates, restir_spatial, restir_mcap;
    float restir_radius;
    uint64_t seed;
};

// Per-pixel structure-of-arrays buffers of the ReSTIR path.
struct GBufferSoA {
    float4* p0;  // pos.xyz, kind (int bits)
    float4* p1;  // normal.xyz, rough
    float4* p2;  // albedo.xyz, metallic
    float4* p3;  // wo.xyz, ior
    float4* p4;  // emission.xyz, -
};
struct ReservoirSoA {
    float4* r0;  // light_pos.xyz, p_hat
    float4* r1;  // light_normal.xyz, alpha
    float4* r2;  // emission.xyz, sum_ws
    float4* r3;  // sum_wu, M (uint bits), valid (uint bits), -
};

}  // namespace rmh
