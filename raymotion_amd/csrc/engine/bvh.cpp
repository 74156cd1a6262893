// Binned-SAH BVH build and bottom-up refit over Scene::tris.
//
// The tree is the one src/engine/bvh.cpp:29-272 of the reference builds: the
// same breadth-first job order, unscaled centroids (v0+v1+v2), 16 bins on each
// of the 3 axes, cost (A_L*N_L + A_R*N_R)/A_P against a leaf cost of N, the same
// partition/sort fallbacks and the same child numbering. Keeping the steps
// alike is what makes node arrays comparable with the reference builder.
#include "raymotion/bvh.hpp"

#include <algorithm>
#include <limits>
#include <numeric>
#include <queue>

namespace raymotion {
namespace {

constexpr double kInf = std::numeric_limits<double>::infinity();
constexpr int kBins = 16;

struct Bounds {
    Vec3 lo{kInf, kInf, kInf};
    Vec3 hi{-kInf, -kInf, -kInf};
    void grow(const Vec3& a, const Vec3& b) {
        if (a.x < lo.x) lo.x = a.x;
        if (a.y < lo.y) lo.y = a.y;
        if (a.z < lo.z) lo.z = a.z;
        if (b.x > hi.x) hi.x = b.x;
        if (b.y > hi.y) hi.y = b.y;
        if (b.z > hi.z) hi.z = b.z;
    }
    void grow(const Vec3& p) { grow(p, p); }
    double surface_area() const {  // 2(xy + yz + zx)
        const double dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    }
};

inline double component(const Vec3& v, int axis) { return axis == 0 ? v.x : (axis == 1 ? v.y : v.z); }

// Sum of the three vertices along one axis (three times the centroid).
inline double centroid3(const Triangle& t, int axis) {
    if (axis == 0) return t.v0.x + t.v1.x + t.v2.x;
    if (axis == 1) return t.v0.y + t.v1.y + t.v2.y;
    return t.v0.z + t.v1.z + t.v2.z;
}

inline int longest_axis(const Bounds& b) {
    const double sx = b.hi.x - b.lo.x, sy = b.hi.y - b.lo.y, sz = b.hi.z - b.lo.z;
    return (sx > sy) ? ((sx > sz) ? 0 : 2) : ((sy > sz) ? 1 : 2);
}

}  // namespace

void build_bvh(Scene& scene, int max_leaf_tris) {
    max_leaf_tris = std::max(1, max_leaf_tris);
    scene.clear_bvh();
    const size_t n = scene.tris.size();
    if (n == 0) return;
    scene.max_leaf_tris = max_leaf_tris;
    scene.bvh_tri.resize(n);
    std::iota(scene.bvh_tri.begin(), scene.bvh_tri.end(), 0);

    struct Job { int offset, count, node; };
    scene.nodes.reserve(2 * n);
    scene.nodes.emplace_back();
    std::queue<Job> pending;
    pending.push({0, int(n), 0});

    auto by_axis = [&](int axis) {
        return [&scene, axis](int a, int b) {
            return centroid3(scene.tris[size_t(a)], axis) < centroid3(scene.tris[size_t(b)], axis);
        };
    };
    auto split_node = [&](const Job& job, int mid) {
        const int left = int(scene.nodes.size());
        scene.nodes.emplace_back();
        const int right = int(scene.nodes.size());
        scene.nodes.emplace_back();
        BvhNode& node = scene.nodes[size_t(job.node)];
        node.left = left;
        node.right = right;
        node.offset = -1;
        node.count = 0;
        pending.push({job.offset, mid, left});
        pending.push({job.offset + mid, job.count - mid, right});
    };

    while (!pending.empty()) {
        const Job job = pending.front();
        pending.pop();
        int* ids = scene.bvh_tri.data() + job.offset;

        Bounds box, cbox;
        for (int k = 0; k < job.count; ++k) {
            const Triangle& t = scene.tris[size_t(ids[k])];
            box.grow(t.mn, t.mx);
            cbox.grow({t.v0.x + t.v1.x + t.v2.x, t.v0.y + t.v1.y + t.v2.y, t.v0.z + t.v1.z + t.v2.z});
        }
        {
            BvhNode& node = scene.nodes[size_t(job.node)];
            node.mn = box.lo;
            node.mx = box.hi;
            if (job.count <= max_leaf_tris) {
                node.offset = job.offset;
                node.count = job.count;
                node.left = node.right = -1;
                continue;
            }
        }

        const double parent_area = box.surface_area();
        if (!(parent_area > 0)) {
            // Zero-area node: order along the longest axis and halve.
            std::sort(ids, ids + job.count, by_axis(longest_axis(box)));
            split_node(job, job.count / 2);
            continue;
        }

        double best_cost = double(job.count);
        int best_axis = -1, best_bin = -1;
        for (int axis = 0; axis < 3; ++axis) {
            const double c0 = component(cbox.lo, axis), c1 = component(cbox.hi, axis);
            const double span = c1 - c0;
            if (!(span > 0)) continue;
            int count[kBins] = {};
            Bounds bins[kBins];
            const double scale = kBins / span;
            for (int k = 0; k < job.count; ++k) {
                const Triangle& t = scene.tris[size_t(ids[k])];
                int b = int((centroid3(t, axis) - c0) * scale);
                b = std::min(std::max(b, 0), kBins - 1);
                ++count[b];
                bins[b].grow(t.mn, t.mx);
            }
            // Suffix sweep: bounds and counts of bins [b, kBins).
            int suffix_count[kBins];
            Bounds suffix[kBins];
            Bounds acc;
            int total = 0;
            for (int b = kBins - 1; b >= 0; --b) {
                total += count[b];
                suffix_count[b] = total;
                if (count[b] > 0) acc.grow(bins[b].lo, bins[b].hi);
                suffix[b] = acc;
            }
            // Prefix sweep evaluates every split plane between bins b and b+1.
            Bounds prefix;
            int prefix_count = 0;
            for (int b = 0; b < kBins - 1; ++b) {
                prefix_count += count[b];
                if (count[b] > 0) prefix.grow(bins[b].lo, bins[b].hi);
                const int rc = suffix_count[b + 1];
                if (prefix_count == 0 || rc == 0) continue;
                const double cost =
                    (prefix.surface_area() * prefix_count + suffix[b + 1].surface_area() * rc) / parent_area;
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = axis;
                    best_bin = b;
                }
            }
        }

        int mid = -1;
        if (best_axis >= 0) {
            const double c0 = component(cbox.lo, best_axis), c1 = component(cbox.hi, best_axis);
            const double plane = c0 + (c1 - c0) * (best_bin + 1) / double(kBins);
            int* cut = std::partition(ids, ids + job.count, [&](int a) {
                return centroid3(scene.tris[size_t(a)], best_axis) < plane;
            });
            mid = int(cut - ids);
            if (mid == 0 || mid == job.count) mid = -1;
        }
        if (mid < 0) {
            std::sort(ids, ids + job.count, by_axis(longest_axis(box)));
            mid = job.count / 2;
        }
        split_node(job, mid);
    }
}

void refit_bvh(Scene& scene) {
    if (scene.bvh_tri.size() != scene.tris.size() || scene.nodes.empty()) {
        build_bvh(scene);
        return;
    }
    // Children always have larger indices than their parent (breadth-first
    // numbering), so a reverse sweep sees every child before its parent.
    for (size_t i = scene.nodes.size(); i-- > 0;) {
        BvhNode& node = scene.nodes[i];
        Vec3 lo{kInf, kInf, kInf}, hi{-kInf, -kInf, -kInf};
        auto merge = [&](const Vec3& a, const Vec3& b) {
            lo = {std::min(lo.x, a.x), std::min(lo.y, a.y), std::min(lo.z, a.z)};
            hi = {std::max(hi.x, b.x), std::max(hi.y, b.y), std::max(hi.z, b.z)};
        };
        if (node.is_leaf()) {
            for (int k = 0; k < node.count; ++k) {
                const Triangle& t = scene.tris[size_t(scene.bvh_tri[size_t(node.offset + k)])];
                merge(t.mn, t.mx);
            }
        } else {
            merge(scene.nodes[size_t(node.left)].mn, scene.nodes[size_t(node.left)].mx);
            merge(scene.nodes[size_t(node.right)].mn, scene.nodes[size_t(node.right)].mx);
        }
        node.mn = lo;
        node.mx = hi;
    }
}

}  // namespace raymotion
