#pragma once
// Camera and render settings (interface of include/raymotion/pathtrace.hpp:12-81
// in the reference, defaults unchanged). denoise_iterations is an addition:
// the number of a-trous passes (the reference hard-codes 2).
#include <cstdint>
#include <functional>
#include <vector>

#include "raymotion/math.hpp"
#include "raymotion/scene.hpp"

namespace raymotion {

struct Camera {
    double x = 0, y = 0, z = 0;
    double dirx = 0, diry = 0, dirz = 0;  // Euler angles in degrees
    double fov = 60;                      // degrees
    double focal = 201.3839114815688;     // recomputed by update_focal()
    double m0 = 1, m1 = 0, m2 = 0;        // row-major camera-to-world rotation
    double m3 = 0, m4 = 1, m5 = 0;
    double m6 = 0, m7 = 0, m8 = 1;
    struct VectorView {
        const double &x, &y, &z;
        operator Vec3() const { return {x, y, z}; }
    };
    struct View {
        VectorView position, rotation;
        const double& fov;
    };
    // Read-only views that always refer to this camera, also after copies.
    const View get{{x, y, z}, {dirx, diry, dirz}, fov};

    Camera() = default;
    Camera(const Camera& other) : Camera() { *this = other; }
    Camera& operator=(const Camera& o) {
        x = o.x; y = o.y; z = o.z;
        dirx = o.dirx; diry = o.diry; dirz = o.dirz;
        fov = o.fov; focal = o.focal;
        m0 = o.m0; m1 = o.m1; m2 = o.m2;
        m3 = o.m3; m4 = o.m4; m5 = o.m5;
        m6 = o.m6; m7 = o.m7; m8 = o.m8;
        return *this;
    }
    void update_trig();   // m0..m8 from dirx/diry/dirz
    void update_focal();  // focal = 240 / tan(fov / 2), independent of resolution
    void set(Vec3 position = {}, Vec3 rotation = {}, double field_of_view = 60) {
        x = position.x; y = position.y; z = position.z;
        dirx = rotation.x; diry = rotation.y; dirz = rotation.z;
        fov = field_of_view;
        update_trig();
        update_focal();
    }
};

struct RenderConfig {
    int width = 640;
    int height = 360;
    int spp = 1;
    int max_bounces = 5;
    double resolution = 1;   // sub-pixel jitter width
    double clamp = 16;       // per-sample radiance clamp; 0 disables
    int adapt_min = 16;      // adaptive sampling: minimum samples (0 disables)
    int adapt_step = 8;
    double adapt_rel = 0.02;
    double adapt_abs = 0.002;
    int restir_candidates = 8;
    int restir_spatial = 4;
    double restir_radius = 16;
    int restir_mcap = 128;
    int denoise_iterations = 2;
    // Traversal tree of the HIP backend: 0 the scene's host-built tree
    // (scene.nodes); 1 a tree built on the GPU and refit there while the
    // triangle count stays (scene.nodes may be empty); 2 the same, rebuilt for
    // this frame. Not in the reference, whose tree is always host-built.
    int gpu_tree = 0;
};

}  // namespace raymotion
