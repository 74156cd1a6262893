// Host-side conversion of a reference-layout scene (rm_scene_desc) into the
// device layouts of rm_layout.h. Plain C++: no HIP calls, so it can be checked
// on a machine without a GPU.
//
// A scene is packed in three independent parts so that a frame which only
// moves vertices (the animated case) re-packs and re-uploads only geometry:
//   tree      4-wide nodes collapsed from the caller's binary tree;
//   geometry  leaf-ordered vertices and the normals (plus the small light
//             table, re-packed with any change);
//   materials deduplicated material table, per-triangle ids, uvs, texels.
// Each part has a content key computed over fixed-size chunks in parallel
// (deterministic for any thread count).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <functional>
#include <new>
#include <utility>
#include <string>
#include <vector>

#include "rm_hip.h"
#include "rm_keys.h"

namespace rmh {

struct F4 {
    float x, y, z, w;
};

// Host buffers of the scene upload live in pinned memory (hipHostMalloc), so
// the copies to the device are direct DMA, asynchronous to the host; resize()
// leaves new elements uninitialised (every packer writes all of them). The
// upload waits for its copies before the buffers are written again.
template <class T>
struct PinnedAllocator {
    using value_type = T;
    PinnedAllocator() = default;
    template <class U>
    PinnedAllocator(const PinnedAllocator<U>&) noexcept {}
    T* allocate(size_t n) {
        void* p = nullptr;
        if (hipHostMalloc(&p, n * sizeof(T), hipHostMallocDefault) != hipSuccess) throw std::bad_alloc();
        return static_cast<T*>(p);
    }
    void deallocate(T* p, size_t) noexcept { (void)hipHostFree(p); }
    template <class U>
    void construct(U* p) noexcept {
        ::new (static_cast<void*>(p)) U;
    }
    template <class U, class... A>
    void construct(U* p, A&&... a) {
        ::new (static_cast<void*>(p)) U(std::forward<A>(a)...);
    }
    friend bool operator==(const PinnedAllocator&, const PinnedAllocator&) { return true; }
    friend bool operator!=(const PinnedAllocator&, const PinnedAllocator&) { return false; }
};
template <class T>
using PinnedVector = std::vector<T, PinnedAllocator<T>>;

// Device records of one triangle record t: its kTriGeoWords vertex words
// ([v0.xyz v1.x] [v1.yz v2.xy] [v2.z id - -], id = the triangle's index,
// stored as int bits), and its kTriNrmWords shading-normal words plus the
// face normal ([n0.xyz n.x] [n1.xyz n.y] [n2.xyz n.z], [nn.xyz -]).
inline void pack_tri_geo(const rm_triangle& t, int32_t id, F4* g) {
    float idf;
    std::memcpy(&idf, &id, 4);
    g[0] = {t.v0.x, t.v0.y, t.v0.z, t.v1.x};
    g[1] = {t.v1.y, t.v1.z, t.v2.x, t.v2.y};
    g[2] = {t.v2.z, idf, 0.0f, 0.0f};
}
inline void pack_tri_nrm(const rm_triangle& t, F4* n, F4* nn) {
    n[0] = {t.n0.x, t.n0.y, t.n0.z, t.n.x};
    n[1] = {t.n1.x, t.n1.y, t.n1.z, t.n.y};
    n[2] = {t.n2.x, t.n2.y, t.n2.z, t.n.z};
    *nn = {t.nn.x, t.nn.y, t.nn.z, 0.0f};
}

struct PackedScene {
    PinnedVector<F4> nodes;        // kNodeWords per 4-wide node, preorder (parents before children)
    std::vector<int32_t> depth;   // tree depth of each 4-wide node
    int32_t stack_need = 0;       // deepest possible traversal stack
    PinnedVector<F4> geo;          // kTriGeoWords per triangle, leaf order
    PinnedVector<F4> nrm;          // kTriNrmWords per triangle, scene order
    PinnedVector<F4> nn;           // 1 per triangle
    PinnedVector<F4> lights;       // kLightWords per light, CDF order
    PinnedVector<float> cdf;
    float light_total = 0.0f;
    PinnedVector<int32_t> matid;   // 1 per triangle
    PinnedVector<F4> mats;         // kMatWords per material
    PinnedVector<F4> uv;           // 2 per triangle when any triangle is textured, else empty
    PinnedVector<float> texels;
    int32_t opaque = 1;
};

// Content keys. `positions` covers bytes 0..35 of every triangle record (the
// vertices), `normals` bytes 36..95, `material` bytes 96..175 (everything else)
// and the texels, `tree` the node array and bvh_tri, `lights` light_tri, the
// CDF and the light total. Equal bytes give equal keys.
struct SceneKeys {
    uint64_t positions = 0, normals = 0, material = 0, tree = 0, lights = 0;
    uint64_t ntri = 0, nnode = 0, nlight = 0, ntexel = 0;
    bool operator==(const SceneKeys& o) const {
        return positions == o.positions && normals == o.normals && material == o.material && tree == o.tree &&
               lights == o.lights && ntri == o.ntri && nnode == o.nnode && nlight == o.nlight && ntexel == o.ntexel;
    }
};
SceneKeys scene_keys(const rm_scene_desc& s);
// The same keys from the records' per-chunk values (taken by the marshal
// while it wrote s.tris); rk.chunks must be rmkeys::key_chunks(s.ntri).
SceneKeys scene_keys_from_chunks(const rm_scene_desc& s, const rmkeys::RecordKeys& rk);

// Temporal-history key: the whole triangle records, the triangle and light
// counts and the light total (the fields metal.mm:241-248 hashes).
uint64_t history_key(const SceneKeys& k, float light_total);

// Null arrays and counts (always), tree links and leaf ranges (`tree`),
// texture ranges (`materials`), light indices (`lights`). Returns false with a
// message in `err` when the scene is malformed.
bool validate_scene(const rm_scene_desc& s, bool tree, bool materials, bool lights, std::string& err);

void pack_tree(const rm_scene_desc& s, PackedScene& out);       // nodes, depth, stack_need
void pack_geometry(const rm_scene_desc& s, PackedScene& out);   // geo (leaf order: needs the tree's bvh_tri), nrm, nn
// pack_geometry in `slices` consecutive triangle ranges, calling
// sink(first, count) once the geo, nrm and nn entries of triangles
// [first, first + count) are written (the caller copies them to the device
// while the next range is packed). False if sink returned false.
bool pack_geometry_sliced(const rm_scene_desc& s, PackedScene& out, size_t slices,
                          const std::function<bool(size_t, size_t)>& sink);
void pack_lights(const rm_scene_desc& s, PackedScene& out);     // lights, cdf, light_total
void pack_materials(const rm_scene_desc& s, PackedScene& out);  // matid, mats, uv, texels, opaque

// Same triangle count and identical tree topology (escape links, leaf ranges
// and triangle order), i.e. the resident 4-wide tree can be refit in place.
bool same_topology(const rm_scene_desc& s, const std::vector<int32_t>& prev_bvh_tri,
                   const std::vector<int32_t>& prev_topo);
void topology_of(const rm_scene_desc& s, std::vector<int32_t>& bvh_tri, std::vector<int32_t>& topo);

}  // namespace rmh
