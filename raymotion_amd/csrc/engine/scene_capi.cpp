// C-ABI over the host scene engine (include/rm_scene.h).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "marshal.hpp"
#include "rm_parallel.h"
#include "raymotion/bvh.hpp"
#include "raymotion/renderer.hpp"
#include "rm_scene.h"

struct rm_scene {
    raymotion::Scene scene;
    // Records of the last marshal, reused by the next one (an animated scene
    // re-marshals into the same buffers). A handle is used by one thread at a time.
    mutable raymotion::MarshalledScene marshalled;
};

using raymotion::Vec3;

extern "C" {

rm_scene* rm_scene_create(void) { return new rm_scene(); }
void rm_scene_destroy(rm_scene* s) { delete s; }
void rm_scene_clear(rm_scene* s) {
    if (!s) return;
    s->scene.clear_triangles();
    s->scene.clear_bvh();
}

int rm_scene_add_triangles(rm_scene* s, size_t n, const double* pos, const double* normals, const double* uv,
                           const double* mat, int32_t* index_out) {
    if (!s || (n > 0 && (!pos || !mat))) return RM_ERR_INVALID;
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
            n0 = n1 = n2 = raymotion::normalize(raymotion::cross(b - a, c - a));
        }
        const double zero[6] = {0, 0, 0, 0, 0, 0};
        const double* t = uv ? uv + 6 * i : zero;
        const double* m = mat + 11 * i;
        const int idx = s->scene.add_triangle(a, b, c, t[0], t[1], t[2], t[3], t[4], t[5], n0, n1, n2, m[0], m[1],
                                              m[2], m[3], m[4], m[5], m[6], m[7], m[8], int(m[10]), m[9]);
        if (index_out) index_out[i] = idx;
    }
    return RM_OK;
}

int rm_scene_set_positions(rm_scene* s, size_t n, const double* pos, const double* normals) {
    if (!s || n != s->scene.tris.size() || (n > 0 && !pos)) return RM_ERR_INVALID;
    raymotion::Scene& sc = s->scene;
    constexpr size_t kChunk = size_t(1) << 14;
    const size_t chunks = (n + kChunk - 1) / kChunk;
    // Reject the update before touching the scene if a triangle would become
    // degenerate (the test of Scene::add_triangle), so the count stays.
    std::vector<char> bad(chunks, 0);
    rmpar::parallel_chunks(chunks, [&](size_t ch) {
        const size_t a0 = ch * kChunk, a1 = std::min(n, a0 + kChunk);
        for (size_t i = a0; i < a1 && !bad[ch]; ++i) {
            const double* p = pos + 9 * i;
            const Vec3 e1 = Vec3{p[3], p[4], p[5]} - Vec3{p[0], p[1], p[2]};
            const Vec3 e2 = Vec3{p[6], p[7], p[8]} - Vec3{p[0], p[1], p[2]};
            const double g00 = dot(e1, e1), g01 = dot(e1, e2), g11 = dot(e2, e2);
            if (std::abs(g00 * g11 - g01 * g01) <= 1e-12) bad[ch] = 1;
        }
    });
    for (char b : bad)
        if (b) return RM_ERR_INVALID;
    // New precompute in place (add_triangle's arithmetic), then the lights in
    // triangle order.
    rmpar::parallel_chunks(chunks, [&](size_t ch) {
        const size_t a0 = ch * kChunk, a1 = std::min(n, a0 + kChunk);
        for (size_t i = a0; i < a1; ++i) {
            raymotion::Triangle& t = sc.tris[i];
            const double* p = pos + 9 * i;
            double gram = 0;
            (void)raymotion::Scene::precompute_triangle({p[0], p[1], p[2]}, {p[3], p[4], p[5]}, {p[6], p[7], p[8]}, t,
                                                        gram);
            if (normals) {
                const double* q = normals + 9 * i;
                t.n0 = {q[0], q[1], q[2]};
                t.n1 = {q[3], q[4], q[5]};
                t.n2 = {q[6], q[7], q[8]};
            }
        }
    });
    sc.rebuild_lights();
    return RM_OK;
}

int rm_scene_build_bvh(rm_scene* s, int max_leaf_tris) {
    if (!s) return RM_ERR_INVALID;
    raymotion::build_bvh(s->scene, max_leaf_tris);
    return RM_OK;
}

int rm_scene_refit_bvh(rm_scene* s) {
    if (!s) return RM_ERR_INVALID;
    raymotion::refit_bvh(s->scene);
    return RM_OK;
}

void rm_scene_counts(const rm_scene* s, size_t* ntri, size_t* nnode, size_t* nlight) {
    if (ntri) *ntri = s ? s->scene.tris.size() : 0;
    if (nnode) *nnode = s ? s->scene.nodes.size() : 0;
    if (nlight) *nlight = s ? s->scene.light_tri.size() : 0;
}

int rm_scene_nodes(const rm_scene* s, double* boxes, int32_t* links, int32_t* bvh_tri) {
    if (!s) return RM_ERR_INVALID;
    const auto& nodes = s->scene.nodes;
    for (size_t i = 0; i < nodes.size(); ++i) {
        const auto& n = nodes[i];
        if (boxes) {
            double* b = boxes + 6 * i;
            b[0] = n.mn.x; b[1] = n.mn.y; b[2] = n.mn.z;
            b[3] = n.mx.x; b[4] = n.mx.y; b[5] = n.mx.z;
        }
        if (links) {
            int32_t* l = links + 4 * i;
            l[0] = n.left; l[1] = n.right; l[2] = n.offset; l[3] = n.count;
        }
    }
    if (bvh_tri && !s->scene.bvh_tri.empty())
        std::memcpy(bvh_tri, s->scene.bvh_tri.data(), s->scene.bvh_tri.size() * sizeof(int32_t));
    return RM_OK;
}

int rm_scene_precompute(const rm_scene* s, double* n, double* nn, double* denom, double* light_cdf,
                        double* light_total) {
    if (!s) return RM_ERR_INVALID;
    const auto& tris = s->scene.tris;
    for (size_t i = 0; i < tris.size(); ++i) {
        if (n) { n[3 * i] = tris[i].n.x; n[3 * i + 1] = tris[i].n.y; n[3 * i + 2] = tris[i].n.z; }
        if (nn) { nn[3 * i] = tris[i].nn.x; nn[3 * i + 1] = tris[i].nn.y; nn[3 * i + 2] = tris[i].nn.z; }
        if (denom) denom[i] = tris[i].denom;
    }
    if (light_cdf && !s->scene.light_cdf.empty())
        std::memcpy(light_cdf, s->scene.light_cdf.data(), s->scene.light_cdf.size() * sizeof(double));
    if (light_total) *light_total = s->scene.light_total;
    return RM_OK;
}

int rm_scene_export(const rm_scene* s, rm_triangle* tris, rm_bvh_node* nodes, int32_t* bvh_tri, int32_t* light_tri,
                    float* light_cdf, float* light_total) {
    if (!s) return RM_ERR_INVALID;
    raymotion::MarshalledScene& m = s->marshalled;
    raymotion::marshal_scene(s->scene, m);
    if (tris && !m.tris.empty()) std::memcpy(tris, m.tris.data(), m.tris.size() * sizeof(rm_triangle));
    if (nodes && !m.nodes.empty()) std::memcpy(nodes, m.nodes.data(), m.nodes.size() * sizeof(rm_bvh_node));
    if (bvh_tri && !s->scene.bvh_tri.empty())
        std::memcpy(bvh_tri, s->scene.bvh_tri.data(), s->scene.bvh_tri.size() * sizeof(int32_t));
    if (light_tri && !s->scene.light_tri.empty())
        std::memcpy(light_tri, s->scene.light_tri.data(), s->scene.light_tri.size() * sizeof(int32_t));
    if (light_cdf && !m.cdf.empty()) std::memcpy(light_cdf, m.cdf.data(), m.cdf.size() * sizeof(float));
    if (light_total) *light_total = float(s->scene.light_total);
    return RM_OK;
}

int rm_scene_upload(rm_ctx* ctx, const rm_scene* s) {
    if (!ctx || !s) return RM_ERR_INVALID;
    return raymotion::marshal_and_upload(ctx, s->scene, s->marshalled, false, 0);
}

int rm_scene_upload_gpu_tree(rm_ctx* ctx, const rm_scene* s, int rebuild) {
    if (!ctx || !s) return RM_ERR_INVALID;
    return raymotion::marshal_and_upload(ctx, s->scene, s->marshalled, true, rebuild);
}

static rm_camera to_c(const raymotion::Camera& c) {
    return {float(c.x), float(c.y), float(c.z), float(c.focal), float(c.m0), float(c.m1), float(c.m2),
            float(c.m3), float(c.m4), float(c.m5), float(c.m6), float(c.m7), float(c.m8)};
}

void rm_camera_set(const double pos[3], const double rot_deg[3], double fov_deg, rm_camera* out) {
    raymotion::Camera c;
    c.set({pos[0], pos[1], pos[2]}, {rot_deg[0], rot_deg[1], rot_deg[2]}, fov_deg);
    *out = to_c(c);
}

void rm_camera_orbit(const double center[3], double radius, double height, double angle_deg, double fov_deg,
                     rm_camera* out) {
    raymotion::Camera base;
    base.fov = fov_deg;
    *out = to_c(raymotion::orbit_camera(base, {center[0], center[1], center[2]}, radius, height, angle_deg));
}

void rm_default_frame(rm_frame* out, int width, int height, int spp, int bounces) {
    if (!out) return;
    raymotion::RenderConfig cfg;
    cfg.width = width;
    cfg.height = height;
    if (spp > 0) cfg.spp = spp;
    if (bounces >= 0) cfg.max_bounces = bounces;
    raymotion::Camera cam;
    cam.update_focal();
    cam.update_trig();
    *out = raymotion::make_frame(cam, cfg);
}

}  // extern "C"
