// Scene-program runtime semantics without a GPU: a recording Renderer
// replaces renderer.cpp at link time (the seam tests/test_runtime.cpp of the
// reference uses) and captures the world triangles of each frame.
#include <cmath>
#include <iostream>
#include <stdexcept>

#include "raymotion/runtime.hpp"

namespace {
int builds = 0, frames = 0, gpu_tree = -1;
bool tree_empty = false;
std::vector<raymotion::Triangle> captured;
void expect(bool ok, const char* what) {
    if (!ok) throw std::runtime_error(std::string("runtime check failed: ") + what);
}
}  // namespace

namespace raymotion {
Renderer::Renderer() = default;
void Renderer::configure(int w, int h, int spp, int b) {
    config.width = w; config.height = h; config.spp = spp; config.max_bounces = b;
}
void Renderer::build(int leaf) { ++builds; build_bvh(scene, leaf); }
std::vector<uint8_t> Renderer::render(uint64_t, ProgressFn) {
    captured = scene.tris;
    gpu_tree = config.gpu_tree;
    tree_empty = scene.nodes.empty();
    ++frames;
    return std::vector<uint8_t>(3);
}
bool Renderer::render_to_ppm(const std::string&, uint64_t, ProgressFn) { return true; }
bool Renderer::render_to_png(const std::string&, uint64_t, ProgressFn) { return true; }
}  // namespace raymotion

int main() {
    using namespace raymotion;
    setenv("RAYMOTION_GPU_BVH", "0", 1);  // the reference's host-tree policy first
    auto model = std::make_shared<Scene>();
    model->add_triangle({0, 0, 0}, {1, 0, 0}, {0, 1, 0}, 0, 0, 0, 0, 0, 0, {0, 0, 1}, {0, 0, 1}, {0, 0, 1}, 1, 1, 1, 0,
                        0, 0, 0, 1, 0.5, 0);
    Objects object([](const auto&, int, int, int) {}, 1, 1, 1);
    object.push(model, {0, 0, 3});
    object.render();
    expect(builds == 1 && captured.size() == 1 && captured[0].v0.z == 3, "first frame builds the world");
    object.camera.x = 2;  // camera motion alone keeps the geometry and the tree
    object.push(model, {0, 0, 3});
    object.render();
    expect(builds == 1 && frames == 2 && object.u_timer == 2.0 / 30, "camera-only frame reuses the tree");
    object.push(model, {0, 0, 4}, {}, {1, 1, 1}, {0.2, 0.3, 0.4});
    object.render();
    expect(captured[0].v0.z == 4 && captured[0].ar == 0.2, "moved + recoloured instance");
    object.render();
    expect(captured.empty(), "push queue is cleared after render");
    object.push(model, {0, 0, 3});
    object.push(model, {0, 0, 5});
    object.render();
    expect(captured.size() == 2 && captured[1].v0.z == 5, "instance order");
    object.push(model, {0, 0, 5});
    object.push(model, {0, 0, 3});
    object.render();
    expect(captured[0].v0.z == 5 && captured[1].v0.z == 3, "reordered instances");
    object.push(model, {}, {0, 0, 90}, {2, 1, 1});
    object.render();
    expect(std::abs(captured[0].v1.y - 2) < 1e-12 && std::abs(captured[0].v1.x) < 1e-12, "rotation and scale");
    // Six geometry updates so far. With unchanged topology the tree is refit,
    // except when the update counter is a multiple of 32 (runtime.hpp:124):
    // updates 6..69 contain two such values, 32 and 64.
    const int before = builds;
    for (int i = 0; i < 64; ++i) {
        object.push(model, {0, 0, 3.0 + i});
        object.render();
    }
    expect(builds - before == 2, "refit, with a rebuild every 32nd geometry update");
    expect(gpu_tree == 0, "host tree when RAYMOTION_GPU_BVH=0");
    // GPU tree (the default): the same policy, carried by config.gpu_tree.
    setenv("RAYMOTION_GPU_BVH", "1", 1);
    Objects gpu([](const auto&, int, int, int) {}, 1, 1, 1);
    const int builds_before = builds;
    gpu.push(model, {0, 0, 3});
    gpu.render();
    expect(gpu_tree == 2 && tree_empty, "first frame: GPU build, no host tree");
    gpu.push(model, {0, 0, 3});
    gpu.render();
    expect(gpu_tree == 1, "camera-only frame keeps the GPU tree");
    int rebuilds = 0;
    for (int i = 0; i < 64; ++i) {
        gpu.push(model, {0, 0, 4.0 + i});  // geometry updates 1..64: rebuilds at 32 and 64
        gpu.render();
        rebuilds += gpu_tree == 2;
        expect(gpu_tree == 1 || gpu_tree == 2, "GPU tree every frame");
    }
    expect(rebuilds == 2 && builds == builds_before, "GPU refit, rebuilt every 32nd update, no host build");
    bool threw = false;
    try { object.push(nullptr); } catch (const std::runtime_error&) { threw = true; }
    expect(threw, "null model rejected");
    threw = false;
    try { object.push(model, {}, {}, {0, 1, 1}); } catch (const std::runtime_error&) { threw = true; }
    expect(threw, "zero scale rejected");
    std::cout << "runtime checks passed\n";
    return 0;
}
