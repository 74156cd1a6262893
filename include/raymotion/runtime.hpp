#pragma once
// Scene-program runtime: the object.init / push / render API that generated
// main.cpp files call (interface of include/raymotion/runtime.hpp:12-148 in the
// reference). World geometry is rebuilt only when the pushed instances
// change; with unchanged topology the BVH is refit, and every 32nd geometry
// update is a full rebuild (runtime.hpp:119-130), on the GPU unless
// RAYMOTION_GPU_BVH=0. Frame seeds are 12345+frame.
#include <cstdlib>
#include <filesystem>
#include <memory>
#include <optional>
#include <stdexcept>
#include <tuple>
#include <unordered_map>

#include "raymotion/bvh.hpp"
#include "raymotion/obj.hpp"
#include "raymotion/renderer.hpp"

namespace raymotion {

struct RenderComplete {};
struct Resolution { double x, y; };
using FrameSink = std::function<void(const std::vector<uint8_t>&, int, int, int)>;
using Model = std::shared_ptr<const Scene>;

class Objects {
    struct Instance {
        Model model;
        Vec3 position, rotation, scale, albedo, emission;
        std::optional<double> refract, rougth, metallic, alpha;
        auto key() const {
            return std::tie(model, position.x, position.y, position.z, rotation.x, rotation.y, rotation.z, scale.x,
                            scale.y, scale.z, albedo.x, albedo.y, albedo.z, emission.x, emission.y, emission.z,
                            refract, rougth, metallic, alpha);
        }
        bool operator==(const Instance& o) const { return key() == o.key(); }
    };
    std::vector<Instance> instances, previous_instances;
    std::vector<const Scene*> topology, previous_topology;
    std::unordered_map<std::string, Model> models;
    bool scene_ready = false;
    unsigned geometry_updates = 0;
    Renderer renderer;
    FrameSink frame_sink;
    std::filesystem::path output;
    bool video;
    int frame = 0;
    int framerate;
    int max_frames;
    double timer = 0;

    // RAYMOTION_GPU_BVH=0 keeps the reference's host-built tree; otherwise the
    // HIP backend builds and refits the tree on the GPU (same images).
    static bool gpu_bvh() {
        const char* v = std::getenv("RAYMOTION_GPU_BVH");
        return !(v && v[0] == '0' && v[1] == '\0');
    }

    void append(const Instance& in) {
        topology.push_back(in.model.get());
        const double cx = cos_deg(in.rotation.x), sx = sin_deg(in.rotation.x);
        const double cy = cos_deg(in.rotation.y), sy = sin_deg(in.rotation.y);
        const double cz = cos_deg(in.rotation.z), sz = sin_deg(in.rotation.z);
        // Rotation order x, then y, then z.
        auto rotate = [&](Vec3 v) {
            v = {v.x, cx * v.y - sx * v.z, sx * v.y + cx * v.z};
            v = {cy * v.x + sy * v.z, v.y, -sy * v.x + cy * v.z};
            return Vec3{cz * v.x - sz * v.y, sz * v.x + cz * v.y, v.z};
        };
        const Vec3 s = in.scale;
        auto point = [&](const Vec3& v) { return rotate({v.x * s.x, v.y * s.y, v.z * s.z}) + in.position; };
        auto normal = [&](const Vec3& v) { return normalize(rotate({v.x / s.x, v.y / s.y, v.z / s.z})); };
        Scene& world = renderer.scene;
        world.tris.reserve(world.tris.size() + in.model->tris.size());
        for (const Triangle& t : in.model->tris) {
            const int index = world.add_triangle(
                point(t.v0), point(t.v1), point(t.v2), t.tu0, t.tv0, t.tu1, t.tv1, t.tu2, t.tv2, normal(t.n0),
                normal(t.n1), normal(t.n2), in.albedo.x < 0 ? t.ar : in.albedo.x, in.albedo.y < 0 ? t.ag : in.albedo.y,
                in.albedo.z < 0 ? t.ab : in.albedo.z, in.emission.x < 0 ? t.er : in.emission.x,
                in.emission.y < 0 ? t.eg : in.emission.y, in.emission.z < 0 ? t.eb : in.emission.z,
                in.metallic.value_or(t.metallic), in.refract.value_or(t.ior), in.rougth.value_or(t.rough), t.shader,
                in.alpha.value_or(t.alpha));
            if (index >= 0) world.tris[size_t(index)].texture = t.texture;
        }
    }

public:
    Camera& camera;
    const double& u_timer = timer;
    const Resolution u_resolution;

    Objects(std::string path, int width, int height, int sample, bool movie, int fps = 30, int frame_limit = 0)
        : output(std::move(path)), video(movie), framerate(fps), max_frames(frame_limit), camera(renderer.camera),
          u_resolution{double(width), double(height)} {
        if (fps <= 0) throw std::invalid_argument("framerate must be positive");
        if (frame_limit < 0) throw std::invalid_argument("frame limit must be nonnegative");
        renderer.configure(width, height, sample);
    }
    // A synchronous sink receives every frame (backpressure); no files are written.
    Objects(FrameSink sink, int width, int height, int sample, int fps = 30)
        : Objects("", width, height, sample, true, fps) {
        if (!sink) throw std::invalid_argument("frame sink is required");
        frame_sink = std::move(sink);
    }

    Model init(const std::string& path, const std::string& mtl = "") {
        const std::string key = path + std::string(1, '\0') + mtl;
        if (auto it = models.find(key); it != models.end()) return it->second;
        auto model = std::make_shared<Scene>();
        const int added = mtl.empty() ? add_obj_file(*model, path, 0, 0, 0, 1, 1, 1, 1, 0, 0, 0, 1, 0.5, 0)
                                      : load_obj_materials(*model, path, mtl);
        if (added < 0) throw std::runtime_error("cannot load OBJ: " + path);
        if (model->tris.empty()) throw std::runtime_error("OBJ has no valid triangles: " + path);
        models[key] = model;
        return model;
    }

    void push(const Model& model, Vec3 position = {}, Vec3 rotation = {}, Vec3 scale = {1, 1, 1},
              Vec3 albedo = {-1, -1, -1}, Vec3 emission = {-1, -1, -1}, std::optional<double> refract = {},
              std::optional<double> rougth = {}, std::optional<double> metallic = {},
              std::optional<double> alpha = {}) {
        if (!model) throw std::runtime_error("null object");
        if (scale.x == 0 || scale.y == 0 || scale.z == 0) throw std::runtime_error("scale must be nonzero");
        instances.push_back({model, position, rotation, scale, albedo, emission, refract, rougth, metallic, alpha});
    }

    void render() {
        if (video && max_frames > 0 && frame >= max_frames) throw RenderComplete{};
        if (!video && frame > 0) throw std::runtime_error("PNG requires exactly one object.render()");
        if (!scene_ready || instances != previous_instances) {
            renderer.scene.clear_triangles();
            topology.clear();
            for (const Instance& i : instances) append(i);
            const bool rebuild = !(topology == previous_topology && geometry_updates % 32 != 0);
            if (gpu_bvh()) {
                // The backend builds (rebuild) or refits the tree on the GPU.
                renderer.scene.clear_bvh();
                renderer.config.gpu_tree = rebuild ? 2 : 1;
            } else {
                renderer.config.gpu_tree = 0;
                if (rebuild) renderer.build();
                else refit_bvh(renderer.scene);
            }
            ++geometry_updates;
            previous_topology = topology;
            previous_instances = instances;
            scene_ready = true;
        }
        camera.update_trig();
        camera.update_focal();
        bool ok;
        if (frame_sink) {
            const auto pixels = renderer.render(12345 + uint64_t(frame));
            frame_sink(pixels, renderer.config.width, renderer.config.height, frame);
            ok = true;
        } else if (video) {
            char name[40];
            std::snprintf(name, sizeof(name), "f%08d.ppm", frame);
            ok = renderer.render_to_ppm((output / name).string(), 12345 + uint64_t(frame));
        } else {
            ok = renderer.render_to_png(output.string());
        }
        if (!ok) throw std::runtime_error("cannot write rendered frame");
        if (renderer.config.gpu_tree == 2) renderer.config.gpu_tree = 1;  // rebuilt once
        ++frame;
        timer = double(frame) / framerate;
        instances.clear();
        if (video && max_frames > 0 && frame >= max_frames) throw RenderComplete{};
    }

    void finish() {
        if (frame == 0) throw std::runtime_error("no object.render() executed");
    }
};

}  // namespace raymotion
