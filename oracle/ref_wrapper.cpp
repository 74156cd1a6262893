// TEST INFRASTRUCTURE ONLY. C entry points over the reference's own CPU scene
// engine (scene.cpp, bvh.cpp, camera.cpp compiled from /root/reference by
// oracle/Makefile into oracle/_ref/libref.so). Used by tests to check the
// product engine's triangle precompute, BVH build/refit and camera matrices
// against the reference implementation bit for bit.
#include <cstdint>
#include <cstring>

#include "raymotion/bvh.hpp"
#include "raymotion/pathtrace.hpp"

using namespace raymotion;

struct ref_scene {
    Scene scene;
};

extern "C" {

ref_scene* ref_scene_create(void) { return new ref_scene(); }
void ref_scene_destroy(ref_scene* s) { delete s; }
void ref_scene_clear(ref_scene* s) { s->scene.clear_triangles(); }

// Same argument layout as rm_scene_add_triangles.
void ref_add_triangles(ref_scene* s, size_t n, const double* pos, const double* normals, const double* uv,
                       const double* mat, int32_t* index_out) {
    for (size_t i = 0; i < n; ++i) {
        const double* p = pos + 9 * i;
        const Vec3 a{p[0], p[1], p[2]}, b{p[3], p[4], p[5]}, c{p[6], p[7], p[8]};
        Vec3 n0, n1, n2;
        if (normals) {
            const double* q = normals + 9 * i;
            n0 = {q[0], q[1], q[2]};
            n1 = {q[3], q[4], q[5]};
            n2 = {q[6], q[7], q[8]};
        } else {
            n0 = n1 = n2 = normalize(cross(b - a, c - a));
        }
        const double zero[6] = {0, 0, 0, 0, 0, 0};
        const double* t = uv ? uv + 6 * i : zero;
        const double* m = mat + 11 * i;
        const int idx = s->scene.add_triangle(a, b, c, t[0], t[1], t[2], t[3], t[4], t[5], n0, n1, n2, m[0], m[1],
                                              m[2], m[3], m[4], m[5], m[6], m[7], m[8], int(m[10]), m[9]);
        if (index_out) index_out[i] = idx;
    }
}

void ref_build_bvh(ref_scene* s, int leaf) { build_bvh(s->scene, leaf); }
void ref_refit_bvh(ref_scene* s) { refit_bvh(s->scene); }

void ref_counts(const ref_scene* s, size_t* ntri, size_t* nnode, size_t* nlight) {
    *ntri = s->scene.tris.size();
    *nnode = s->scene.nodes.size();
    *nlight = s->scene.light_tri.size();
}

void ref_nodes(const ref_scene* s, double* boxes, int32_t* links, int32_t* bvh_tri) {
    const auto& nodes = s->scene.nodes;
    for (size_t i = 0; i < nodes.size(); ++i) {
        const auto& n = nodes[i];
        double* b = boxes + 6 * i;
        b[0] = n.mn.x; b[1] = n.mn.y; b[2] = n.mn.z;
        b[3] = n.mx.x; b[4] = n.mx.y; b[5] = n.mx.z;
        int32_t* l = links + 4 * i;
        l[0] = n.left; l[1] = n.right; l[2] = n.offset; l[3] = n.count;
    }
    if (!s->scene.bvh_tri.empty())
        std::memcpy(bvh_tri, s->scene.bvh_tri.data(), s->scene.bvh_tri.size() * sizeof(int32_t));
}

void ref_precompute(const ref_scene* s, double* n, double* nn, double* denom, double* light_cdf, double* light_total) {
    const auto& tris = s->scene.tris;
    for (size_t i = 0; i < tris.size(); ++i) {
        n[3 * i] = tris[i].n.x; n[3 * i + 1] = tris[i].n.y; n[3 * i + 2] = tris[i].n.z;
        nn[3 * i] = tris[i].nn.x; nn[3 * i + 1] = tris[i].nn.y; nn[3 * i + 2] = tris[i].nn.z;
        denom[i] = tris[i].denom;
    }
    if (!s->scene.light_cdf.empty())
        std::memcpy(light_cdf, s->scene.light_cdf.data(), s->scene.light_cdf.size() * sizeof(double));
    *light_total = s->scene.light_total;
}

// Camera::update_trig / update_focal: out = x y z focal m0..m8.
void ref_camera(const double* pos, const double* rot, double fov, double* out) {
    Camera c;
    c.set({pos[0], pos[1], pos[2]}, {rot[0], rot[1], rot[2]}, fov);
    const double v[13] = {c.x, c.y, c.z, c.focal, c.m0, c.m1, c.m2, c.m3, c.m4, c.m5, c.m6, c.m7, c.m8};
    std::memcpy(out, v, sizeof(v));
}

}  // extern "C"
