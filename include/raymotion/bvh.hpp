#pragma once
// CPU SAH BVH over Scene::tris (interface of include/raymotion/bvh.hpp:13-14 in
// the reference). The HIP backend collapses this binary tree into 4-wide nodes.
#include "raymotion/scene.hpp"

namespace raymotion {

// Recomputes node bounds bottom-up; rebuilds (leaf size 8) when the triangle
// count no longer matches the tree or no tree exists.
void refit_bvh(Scene& scene);
// Binned-SAH build (16 bins, 3 axes), breadth-first, leaves of <= max_leaf_tris.
void build_bvh(Scene& scene, int max_leaf_tris = 8);

}  // namespace raymotion
