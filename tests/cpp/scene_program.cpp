// A scene program written against the kept API (the shape of the main() that
// the .ray compiler generates): Objects constructed from argv, init/push/render
// per frame, u_timer and u_resolution, camera edits. Renders through HIP.
#include <raymotion/runtime.hpp>
#include <cmath>
#include <iostream>

using namespace raymotion;

int main(int argc, char** argv) {
    try {
        if (argc != 8) throw std::runtime_error("usage: out w h spp fps png|mp4 frames");
        const bool video = std::string(argv[6]) == "mp4";
        Objects object(argv[1], std::stoi(argv[2]), std::stoi(argv[3]), std::stoi(argv[4]), video,
                       std::stoi(argv[5]), std::stoi(argv[7]));
        auto& camera = object.camera;
        const auto& u_timer = object.u_timer;
        do {
            const double frame_start = u_timer;
            auto tri = std::make_shared<Scene>();
            tri->add_triangle({-20, -20, 3}, {-20, 20, 3}, {20, 20, 3}, 0, 0, 0, 0, 0, 0, {0, 0, -1}, {0, 0, -1},
                              {0, 0, -1}, 1, 1, 1, 1, 0, 0, 0, 1, 0.5, 0);
            tri->add_triangle({-20, -20, 3}, {20, 20, 3}, {20, -20, 3}, 0, 0, 0, 0, 0, 0, {0, 0, -1}, {0, 0, -1},
                              {0, 0, -1}, 1, 1, 1, 1, 0, 0, 0, 1, 0.5, 0);
            camera.fov = 60 + 10 * std::sin(u_timer);
            object.push(tri, {0, 0, 0}, {0, 0, 0}, {1, 1, 1});
            object.render();
            object.finish();
            if (u_timer == frame_start) throw std::runtime_error("no render");
        } while (video);
        return 0;
    } catch (const RenderComplete&) {
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "raymotion: " << e.what() << "\n";
        return 1;
    }
}
