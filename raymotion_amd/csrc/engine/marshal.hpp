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
    rm_scene_desc desc(const Scene& scene) const;
    rmkeys::RecordKeys keys() const;
};

// FP64 triangles -> 176-byte FP32 records with a deduplicated texture atlas
// (metal.mm:109-128), binary BVH -> preorder array with escape links
// (metal.mm:129-144), light CDF -> FP32 (metal.mm:145). `out` may hold the
// previous marshal's arrays: records that did not change are not rewritten.
void marshal_scene(const Scene& scene, MarshalledScene& out);

// Reference host-side parameter handling (metal.mm:146-170, 251-252, 269-271):
// defaults, clamps and environment switches.
rm_frame make_frame(const Camera& camera, const RenderConfig& config);

}  // namespace raymotion
