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

// Tree node with kWide children: 2 * kWide float4 words. With kGroups = kWide / 4
// groups of four children, word node_lo_word(a, g) holds the lower bounds on
// axis a of children 4g .. 4g+3 and node_hi_word(a, g) their upper bounds;
// then kGroups words of child references and kGroups words of child kinds
// (int bits):
//   kind == -1 empty slot, kind == 0 internal child (ref = node index),
//   kind  >  0 leaf with `kind` triangles starting at leaf-order index ref.
constexpr int kWide = 4;  // 8 measured: a third fewer node visits, 5-8 % slower frames
constexpr int kGroups = kWide / 4;
constexpr int kNodeWords = 2 * kWide;
constexpr int node_lo_word(int a, int g) { return a * 2 * kGroups + g; }
constexpr int node_hi_word(int a, int g) { return a * 2 * kGroups + kGroups + g; }
constexpr int kNodeRefWord = 6 * kGroups;
constexpr int kNodeKindWord = 7 * kGroups;
static_assert(kWide == 4 || kWide == 8, "node width");

// Node access for host (packing) and device (build, refit) code; F is a
// 4-float struct (float4 or the packer's F4).
__host__ __device__ inline float rm_ibf(int i) {
    float f;
    __builtin_memcpy(&f, &i, 4);
    return f;
}
__host__ __device__ inline int rm_fbi(float f) {
    int i;
    __builtin_memcpy(&i, &f, 4);
    return i;
}
template <class F>
__host__ __device__ inline float& f4_at(F& v, int k) {
    return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w;
}
template <class F>
__host__ __device__ inline float f4_get(const F& v, int k) {
    return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w;
}
// Writes a node from child boxes lo/hi [axis][child], references and kinds.
template <class F>
__host__ __device__ inline void write_node(F* w, const float lo[3][kWide], const float hi[3][kWide],
                                           const int ref[kWide], const int kind[kWide]) {
    for (int c = 0; c < kWide; ++c) {
        const int g = c / 4, j = c % 4;
        for (int a = 0; a < 3; ++a) {
            f4_at(w[node_lo_word(a, g)], j) = lo[a][c];
            f4_at(w[node_hi_word(a, g)], j) = hi[a][c];
        }
        f4_at(w[kNodeRefWord + g], j) = rm_ibf(ref[c]);
        f4_at(w[kNodeKindWord + g], j) = rm_ibf(kind[c]);
    }
}
template <class F>
__host__ __device__ inline void read_node_links(const F* w, int ref[kWide], int kind[kWide]) {
    for (int c = 0; c < kWide; ++c) {
        ref[c] = rm_fbi(f4_get(w[kNodeRefWord + c / 4], c % 4));
        kind[c] = rm_fbi(f4_get(w[kNodeKindWord + c / 4], c % 4));
    }
}
template <class F>
__host__ __device__ inline void set_node_refs(F* w, const int ref[kWide]) {
    for (int c = 0; c < kWide; ++c) f4_at(w[kNodeRefWord + c / 4], c % 4) = rm_ibf(ref[c]);
}
template <class F>
__host__ __device__ inline void read_node_boxes(const F* w, float lo[3][kWide], float hi[3][kWide]) {
    for (int c = 0; c < kWide; ++c)
        for (int a = 0; a < 3; ++a) {
            lo[a][c] = f4_get(w[node_lo_word(a, c / 4)], c % 4);
            hi[a][c] = f4_get(w[node_hi_word(a, c / 4)], c % 4);
        }
}
constexpr int kTriGeoWords = 3;   // [v0.xyz v1.x] [v1.yz v2.xy] [v2.z id - -]
constexpr int kTriNrmWords = 3;   // [n0.xyz n.x] [n1.xyz n.y] [n2.xyz n.z]
constexpr int kMatWords = 4;      // [ar ag ab alpha] [er eg eb ior] [rough metallic shader texoff] [texw texh - -]
constexpr int kLightWords = 4;    // [v0.xyz v1.x] [v1.yz v2.xy] [v2.z nn.xyz] [er eg eb alpha]
constexpr int kLdsLights = 64;    // scenes with at most this many lights sample them from LDS
constexpr int kBlock = 256;       // threads per block of every pixel pass
// Traversal stack: entries per lane kept in LDS, and the capacity of the deep
// kernel instance (LDS entries first, the rest in per-lane scratch memory).
constexpr int kLdsStack = 64;
constexpr int kSpillStack = 256;

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
