#pragma once
// High-level renderer (interface of include/raymotion/renderer.hpp in the
// reference). Frames go through render_image_hip.
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "raymotion/pathtrace.hpp"
#include "raymotion/scene.hpp"

namespace raymotion {

using ProgressFn = std::function<void(int done, int total)>;

struct VideoOptions {
    int frames = 60;
    int fps = 30;
    uint64_t seed = 12345;   // frame f renders with seed + f
    std::string frames_dir;  // default: out.mp4 + ".frames"
    bool keep_frames = false;
};

Camera lerp_camera(const Camera& a, const Camera& b, double t);
// Camera on a horizontal circle around `center`, looking at it.
Camera orbit_camera(const Camera& base, const Vec3& center, double radius, double height, double angle_deg);

class Renderer {
public:
    Scene scene;
    Camera camera;
    RenderConfig config;

    Renderer();
    void build(int max_leaf_tris = 8);
    void configure(int width, int height, int spp = -1, int bounces = -1);
    void set_camera(const Camera& cam) { camera = cam; }

    // Throws std::runtime_error("HIP: " + diagnostic) when the frame fails.
    std::vector<uint8_t> render(uint64_t seed = 12345, ProgressFn progress = {});
    bool render_to_ppm(const std::string& path, uint64_t seed = 12345, ProgressFn progress = {});
    bool render_to_png(const std::string& path, uint64_t seed = 12345, ProgressFn progress = {});

    bool render_sequence(const std::vector<Camera>& cameras, const std::string& out_dir,
                         const std::string& basename = "f", uint64_t seed = 12345, ProgressFn progress = {});
    bool render_lerp_video(const Camera& a, const Camera& b, const std::string& out_mp4, const VideoOptions& opt = {},
                           ProgressFn progress = {});
    bool render_orbit_video(const std::string& out_mp4, const Vec3& center, double radius, double height,
                            double start_deg, double end_deg, const VideoOptions& opt = {}, ProgressFn progress = {});
};

bool encode_mp4(const std::string& frames_dir, const std::string& basename, const std::string& out_mp4, int fps,
                std::string* err = nullptr);
bool have_ffmpeg();
bool save_ppm(const std::string& path, int w, int h, const std::vector<uint8_t>& rgb, std::string* err = nullptr);
bool save_png(const std::string& path, int w, int h, const std::vector<uint8_t>& rgb, std::string* err = nullptr);

}  // namespace raymotion
