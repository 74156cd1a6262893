#include <raymotion/preview.hpp>
#include <iostream>
#include <cmath>
#include <vector>
using namespace raymotion;
int main(int argc, char** argv) {
  try {
    if(argc!=8) throw std::runtime_error("run this program with raymotion export");
    const int width=std::stoi(argv[2]), height=std::stoi(argv[3]);
    const int sample=std::stoi(argv[4]), framerate=std::stoi(argv[5]);
    const bool video=std::string(argv[6])=="mp4";
    PreviewStream stream;
    Objects object(stream.sink(),width,height,sample,framerate);
    auto& camera=object.camera;
    const auto& u_timer=object.u_timer;
    const auto& u_resolution=object.u_resolution;
    do {
    const double frame_start=u_timer;
#line 1 "main.ray"
// A .ray scene program: an emissive quad loaded from an OBJ file fills the
// view, so every pixel of the frame is (127, 0, 0).

auto quad = object.init("quad.obj");
const std::vector<double> emission = {1, 0, 0};
const char* note = "strings keep void x = object.init( untouched";
camera.fov = 60 + 10 * std::sin(u_timer);  /* the fov animates with u_timer */
object.push(quad, Vec3{0, 0, 3}, Vec3{0, 0, 0}, Vec3{20, 20, 1}, Vec3{1, 1, 1},
            Vec3{emission[0], emission[1], emission[2]});
(void)note;
(void)u_resolution;
object.render();

    object.finish();
    if(u_timer==frame_start) throw std::runtime_error("no object.render() executed in this iteration");
    } while(video);
    return 0;
  } catch(const RenderComplete&) { return 0;
  } catch(const std::exception& e) { std::cerr << "raymotion: " << e.what() << "\n"; return 1; }
}
