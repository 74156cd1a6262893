#pragma once
// Scene description consumed by the BVH builder and the HIP backend.
// Field names and meanings follow include/raymotion/scene.hpp:20-96 of the
// reference so that scene programs written against it compile unchanged.
#include <cstddef>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "raymotion/math.hpp"

namespace raymotion {

// Linear-RGB texture, top row first.
struct Texture {
    int width = 0, height = 0;
    std::vector<Vec3> pixels;
    // Nearest texel with wrap-around; v is flipped (v = 0 is the bottom row).
    Vec3 sample(double u, double v) const;
};

struct Triangle {
    std::shared_ptr<const Texture> texture;
    Vec3 v0, v1, v2;
    Vec3 n;          // unnormalised geometric normal (v1-v0) x (v2-v0)
    double d = 0;    // plane offset: -n.v0
    Vec3 U, V;       // barycentric basis vectors
    double denom = 0;  // |e1|^2 |e2|^2 - (e1.e2)^2
    Vec3 nn;         // normalised geometric normal
    double tu0 = 0, tv0 = 0, tu1 = 0, tv1 = 0, tu2 = 0, tv2 = 0;
    Vec3 n0, n1, n2;  // shading (vertex) normals
    Vec3 mn, mx;      // bounding box
    double ar = 1, ag = 1, ab = 1;
    double er = 0, eg = 0, eb = 0;
    double ior = 1;       // > 1.0001 renders as a dielectric
    double alpha = 1;     // coverage: 0 transparent .. 1 opaque
    double rough = 0.5;
    int shader = 0;       // 1 selects the hue-from-azimuth albedo
    double metallic = 0;
};

// Binary BVH node. Internal nodes have count == 0 and two children; leaves
// reference bvh_tri[offset, offset+count).
struct BvhNode {
    Vec3 mn, mx;
    int left = -1;
    int right = -1;
    int offset = -1;
    int count = 0;
    bool is_leaf() const { return count > 0; }
};

struct Scene {
    std::vector<Triangle> tris;
    std::vector<BvhNode> nodes;
    std::vector<int> bvh_tri;
    int max_leaf_tris = 8;
    std::vector<int> light_tri;      // emissive triangle indices
    std::vector<double> light_cdf;   // running sum of area * luminance
    double light_total = 0;

    void clear_triangles() {
        tris.clear();
        light_tri.clear();
        light_cdf.clear();
        light_total = 0;
    }
    void clear_bvh() {
        nodes.clear();
        bvh_tri.clear();
    }
    // Appends a triangle unless it is degenerate (|denom| <= 1e-12); emissive
    // triangles are added to the light CDF. Returns the new index or -1.
    int add_triangle(const Vec3& p0, const Vec3& p1, const Vec3& p2, double u0, double v0,
                     double u1, double v1, double u2, double v2, const Vec3& n0,
                     const Vec3& n1, const Vec3& n2, double ar, double ag, double ab,
                     double er, double eg, double eb, double metallic, double ior,
                     double rough, int shader, double alpha = 1);
    // add_triangle's per-triangle precompute into t (no light bookkeeping);
    // false when the triangle is degenerate. `gram` = |e1 x e2|^2.
    static bool precompute_triangle(const Vec3& p0, const Vec3& p1, const Vec3& p2, Triangle& t, double& gram);
    // Recomputes light_tri, light_cdf and light_total from the triangles in
    // order, with add_triangle's arithmetic.
    void rebuild_lights();
};

}  // namespace raymotion
