/*
 * rm_scene.h — C-ABI of the host scene engine in librm_hip.so.
 *
 * Lets non-C++ callers (the Python package, tests, bench) build scenes with the
 * engine's own triangle precompute and SAH BVH and hand them to the backend:
 *   rm_scene_add_triangles  Scene::add_triangle     (src/engine/scene.cpp:16-87)
 *   rm_scene_build_bvh      build_bvh               (src/engine/bvh.cpp:29-248)
 *   rm_scene_refit_bvh      refit_bvh               (src/engine/bvh.cpp:250-272)
 *   rm_scene_export         the marshalling of      (src/engine/metal.mm:109-145)
 *   rm_scene_upload         marshalling + rm_set_scene
 *   rm_scene_upload_gpu_tree  marshalling + rm_set_scene_gpu_tree (no host tree needed)
 * Return codes are the RM_* codes of rm_hip.h.
 */
#ifndef RM_SCENE_H
#define RM_SCENE_H

#include "rm_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rm_scene rm_scene;

rm_scene* rm_scene_create(void);
void rm_scene_destroy(rm_scene* s);
/* Removes all triangles and lights (Scene::clear_triangles) and the BVH. */
void rm_scene_clear(rm_scene* s);

/* Adds n triangles. pos: 9n doubles (v0,v1,v2); normals: 9n doubles (n0,n1,n2)
 * or NULL for the unit geometric normal; uv: 6n doubles or NULL for zeros;
 * mat: 11n doubles per triangle: ar ag ab er eg eb metallic ior rough alpha
 * shader. index_out (may be NULL) receives each triangle's index, -1 when it
 * was rejected as degenerate. */
int rm_scene_add_triangles(rm_scene* s, size_t n, const double* pos, const double* normals, const double* uv,
                           const double* mat, int32_t* index_out);
/* Replaces the vertices (and, if given, normals) of the existing triangles,
 * recomputing the per-triangle precompute and the light CDF; triangles keep
 * their materials. Fails with RM_ERR_INVALID, leaving the scene unchanged,
 * when n differs from the triangle count or a triangle would become
 * degenerate (which would change the triangle list). */
int rm_scene_set_positions(rm_scene* s, size_t n, const double* pos, const double* normals);

int rm_scene_build_bvh(rm_scene* s, int max_leaf_tris);
int rm_scene_refit_bvh(rm_scene* s);

/* Sizes: triangles, binary BVH nodes, lights. */
void rm_scene_counts(const rm_scene* s, size_t* ntri, size_t* nnode, size_t* nlight);
/* The engine's binary nodes in build order: boxes 6 doubles per node
 * (mn.xyz, mx.xyz), links 4 ints per node (left, right, offset, count), and
 * bvh_tri (ntri ints). Any pointer may be NULL. */
int rm_scene_nodes(const rm_scene* s, double* boxes, int32_t* links, int32_t* bvh_tri);
/* Per-triangle FP64 precompute: 3 doubles each of n and nn, denom, and the
 * light CDF (nlight doubles) with its total. Any pointer may be NULL. */
int rm_scene_precompute(const rm_scene* s, double* n, double* nn, double* denom, double* light_cdf,
                        double* light_total);
/* Marshals into the reference upload layout: tris (ntri), nodes (nnode, in
 * preorder with escape links), bvh_tri (ntri), light_tri (nlight),
 * light_cdf (nlight). Any pointer may be NULL. */
int rm_scene_export(const rm_scene* s, rm_triangle* tris, rm_bvh_node* nodes, int32_t* bvh_tri, int32_t* light_tri,
                    float* light_cdf, float* light_total);
/* Marshals and calls rm_set_scene. */
int rm_scene_upload(rm_ctx* ctx, const rm_scene* s);
/* Marshals (the scene needs no BVH) and calls rm_set_scene_gpu_tree. */
int rm_scene_upload_gpu_tree(rm_ctx* ctx, const rm_scene* s, int rebuild);

/* Camera::set (camera.cpp:3-22): position, Euler angles in degrees and field
 * of view (focal = 240 / tan(fov/2) pixels, independent of resolution). */
void rm_camera_set(const double pos[3], const double rot_deg[3], double fov_deg, rm_camera* out);
/* orbit_camera (renderer.cpp:153-171): on a circle of `radius` around
 * `center` at height `height`, angle in degrees, looking at the centre. */
void rm_camera_orbit(const double center[3], double radius, double height, double angle_deg, double fov_deg,
                     rm_camera* out);

/* Host-side render settings from RenderConfig defaults (pathtrace.hpp:54-81)
 * and the current environment, as the C++ shim applies them. */
void rm_default_frame(rm_frame* out, int width, int height, int spp, int bounces);

#ifdef __cplusplus
}
#endif
#endif /* RM_SCENE_H */
