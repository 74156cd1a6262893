// Scene -> reference-layout arrays handed to the C-ABI (internal header).
#pragma once
#include <cstdint>
#include <memory>
#include <utility>
#include <string>
#include <vector>

#include "raymotion/pathtrace.hpp"
#include "rm_hip.h"
#include "rm_keys.h"
#include "rm_pack.h"

namespace raymotion {

// Allocator whose resize() leaves new elements uninitialised: the marshal
// writes every record, and value-initialising hundreds of MB first (on one
// thread) would cost more than the conversion.
template <class T>
struct NoInitAllocator : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = NoInitAllocator<U>;
    };
    NoInitAllocator() = default;
    template <class U>
    NoInitAllocator(const NoInitAllocator<U>&) noexcept {}
    template <class U>
    void construct(U* p) noexcept {
        ::new (static_cast<void*>(p)) U;
    }
    template <class U, class... A>
    void construct(U* p, A&&... a) {
        ::new (static_cast<void*>(p)) U(std::forward<A>(a)...);
    }
};

struct MarshalledScene {
    std::vector<rm_triangle, NoInitAllocator<rm_triangle>> tris;
    std::vector<rm_bvh_node> nodes;
    std::vector<float> cdf;
    std::vector<float> texels;
    // per-chunk key values of tris (rm_keys.h), taken while marshalling
    std::vector<uint64_t> kp, kn, km;
    // Device-format geometry of tris in scene order (rmh::pack_tri_geo /
    // pack_tri_nrm), written only by a marshal with a GeometrySink; pinned,
    // so the sink's copies from it run asynchronously.
    rmh::PinnedVector<rmh::F4> geo, nrm, nn;
    bool changed = true;  // the last marshal rewrote at least one record (true before the first)
    rm_scene_desc desc(const Scene& scene) const;
    rmkeys::RecordKeys keys() const;
};

// Receives a marshal's device-format geometry chunk by chunk, from the worker
// thread that finished the chunk (so put() may run on several threads at
// once): triangles [first, first + count) at geo + 3 first, nrm + 3 first and
// nn + first, which stay unchanged until the next marshal. False stops
// nothing but is reported by marshal_scene.
struct GeometrySink {
    virtual ~GeometrySink() = default;
    virtual bool put(size_t first, size_t count, const rmh::F4* geo, const rmh::F4* nrm, const rmh::F4* nn) = 0;
};

// FP64 triangles -> 176-byte FP32 records with a deduplicated texture atlas
// (metal.mm:109-128), binary BVH -> preorder array with escape links
// (metal.mm:129-144), light CDF -> FP32 (metal.mm:145). `out` may hold the
// previous marshal's arrays: records that did not change are not rewritten.
// With a sink, each chunk's device-format geometry is also written and handed
// to it; returns false if the sink refused a chunk. Without `with_tree` the
// node array is left empty (for uploads that build their own tree).
bool marshal_scene(const Scene& scene, MarshalledScene& out, GeometrySink* sink = nullptr, bool with_tree = true);

// Marshals `scene` into `m` and uploads it (rm_set_scene_keyed, or
// rm_set_scene_gpu_tree_keyed with `gpu_tree`). For a GPU-tree upload of a
// scene whose previous marshal changed (an animated scene), the geometry
// goes to the device's staging buffers chunk by chunk while the rest is
// marshalled.
int marshal_and_upload(rm_ctx* ctx, const Scene& scene, MarshalledScene& m, bool gpu_tree, int rebuild);

// Reference host-side parameter handling (metal.mm:146-170, 251-252, 269-271):
// defaults, clamps and environment switches.
rm_frame make_frame(const Camera& camera, const RenderConfig& config);

}  // namespace raymotion
