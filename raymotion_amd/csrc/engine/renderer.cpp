// Renderer, image writers and camera helpers (behaviour of
// src/engine/renderer.cpp in the reference; frames render through HIP).
#include "raymotion/renderer.hpp"

#include <array>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>

#include "raymotion/bvh.hpp"
#include "raymotion/hip.hpp"

namespace raymotion {
namespace fs = std::filesystem;

Renderer::Renderer() {
    camera.update_focal();
    camera.update_trig();
}

void Renderer::build(int max_leaf_tris) { build_bvh(scene, max_leaf_tris); }

void Renderer::configure(int width, int height, int spp, int bounces) {
    config.width = width;
    config.height = height;
    if (spp > 0) config.spp = spp;
    if (bounces > 0) config.max_bounces = bounces;
}

std::vector<uint8_t> Renderer::render(uint64_t seed, ProgressFn progress) {
    std::vector<uint8_t> image;
    std::string error;
    if (!render_image_hip(scene, camera, config, seed, image, error)) throw std::runtime_error("HIP: " + error);
    if (progress) progress(config.height, config.height);
    return image;
}

bool Renderer::render_to_ppm(const std::string& path, uint64_t seed, ProgressFn progress) {
    return save_ppm(path, config.width, config.height, render(seed, std::move(progress)));
}

bool Renderer::render_to_png(const std::string& path, uint64_t seed, ProgressFn progress) {
    return save_png(path, config.width, config.height, render(seed, std::move(progress)));
}

bool Renderer::render_sequence(const std::vector<Camera>& cameras, const std::string& out_dir,
                               const std::string& basename, uint64_t seed, ProgressFn progress) {
    std::error_code ec;
    fs::create_directories(out_dir, ec);
    if (ec) return false;
    for (size_t f = 0; f < cameras.size(); ++f) {
        std::vector<uint8_t> img;
        std::string error;
        if (!render_image_hip(scene, cameras[f], config, seed + uint64_t(f), img, error))
            throw std::runtime_error("HIP: " + error);
        char name[64];
        std::snprintf(name, sizeof(name), "%s%04d.ppm", basename.c_str(), int(f));
        std::string err;
        if (!save_ppm((fs::path(out_dir) / name).string(), config.width, config.height, img, &err)) {
            std::cerr << "frame " << f << ": " << err << "\n";
            return false;
        }
        if (progress) progress(int(f) + 1, int(cameras.size()));
    }
    return true;
}

namespace {
bool finish_video(Renderer& r, const std::vector<Camera>& cams, const std::string& out_mp4, const VideoOptions& opt,
                  ProgressFn progress) {
    const std::string dir = opt.frames_dir.empty() ? out_mp4 + ".frames" : opt.frames_dir;
    if (!r.render_sequence(cams, dir, "f", opt.seed, std::move(progress))) return false;
    std::string err;
    if (!encode_mp4(dir, "f", out_mp4, opt.fps, &err)) {
        std::cerr << err << "\n";
        return false;
    }
    if (!opt.keep_frames) {
        std::error_code ec;
        fs::remove_all(dir, ec);
    }
    return true;
}
}  // namespace

bool Renderer::render_lerp_video(const Camera& a, const Camera& b, const std::string& out_mp4, const VideoOptions& opt,
                                 ProgressFn progress) {
    std::vector<Camera> cams;
    for (int f = 0; f < opt.frames; ++f)
        cams.push_back(lerp_camera(a, b, opt.frames == 1 ? 0.0 : double(f) / (opt.frames - 1)));
    return finish_video(*this, cams, out_mp4, opt, std::move(progress));
}

bool Renderer::render_orbit_video(const std::string& out_mp4, const Vec3& center, double radius, double height,
                                  double start_deg, double end_deg, const VideoOptions& opt, ProgressFn progress) {
    std::vector<Camera> cams;
    for (int f = 0; f < opt.frames; ++f) {
        const double t = opt.frames == 1 ? 0.0 : double(f) / (opt.frames - 1);
        cams.push_back(orbit_camera(camera, center, radius, height, start_deg + (end_deg - start_deg) * t));
    }
    return finish_video(*this, cams, out_mp4, opt, std::move(progress));
}

Camera lerp_camera(const Camera& a, const Camera& b, double t) {
    Camera c = a;
    c.x = a.x + (b.x - a.x) * t;
    c.y = a.y + (b.y - a.y) * t;
    c.z = a.z + (b.z - a.z) * t;
    c.dirx = a.dirx + (b.dirx - a.dirx) * t;
    c.diry = a.diry + (b.diry - a.diry) * t;
    c.dirz = a.dirz + (b.dirz - a.dirz) * t;
    c.fov = a.fov + (b.fov - a.fov) * t;
    c.update_focal();
    c.update_trig();
    return c;
}

Camera orbit_camera(const Camera& base, const Vec3& center, double radius, double height, double angle_deg) {
    Camera c = base;
    const double r = deg2rad(angle_deg);
    c.x = center.x + radius * std::cos(r);
    c.z = center.z + radius * std::sin(r);
    c.y = height;
    const double dx = center.x - c.x, dy = center.y - c.y, dz = center.z - c.z;
    c.diry = rad2deg(std::atan2(dx, dz));                         // yaw towards the centre
    c.dirx = -rad2deg(std::atan2(dy, std::sqrt(dx * dx + dz * dz)));  // pitch towards the centre
    c.update_focal();
    c.update_trig();
    return c;
}

bool save_ppm(const std::string& path, int w, int h, const std::vector<uint8_t>& rgb, std::string* err) {
    if (rgb.size() != size_t(w) * size_t(h) * 3) {
        if (err) *err = "bad buffer size";
        return false;
    }
    std::ofstream f(path, std::ios::binary);
    if (!f) {
        if (err) *err = "cannot open " + path;
        return false;
    }
    f << "P6\n" << w << " " << h << "\n255\n";
    f.write(reinterpret_cast<const char*>(rgb.data()), std::streamsize(rgb.size()));
    return bool(f);
}

namespace {
uint32_t crc32(const uint8_t* p, size_t n) {
    static const std::array<uint32_t, 256> table = [] {
        std::array<uint32_t, 256> t{};
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t c = i;
            for (int k = 0; k < 8; ++k) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            t[i] = c;
        }
        return t;
    }();
    uint32_t c = 0xFFFFFFFFu;
    for (size_t i = 0; i < n; ++i) c = table[(c ^ p[i]) & 0xFFu] ^ (c >> 8);
    return c ^ 0xFFFFFFFFu;
}
void be32(std::vector<uint8_t>& o, uint32_t v) {
    for (int s = 24; s >= 0; s -= 8) o.push_back(uint8_t(v >> s));
}
void chunk(std::vector<uint8_t>& png, const char* type, const std::vector<uint8_t>& data) {
    be32(png, uint32_t(data.size()));
    const size_t start = png.size();
    png.insert(png.end(), type, type + 4);
    png.insert(png.end(), data.begin(), data.end());
    be32(png, crc32(png.data() + start, data.size() + 4));
}
}  // namespace

// 8-bit RGB PNG with stored (uncompressed) deflate blocks; no dependencies.
bool save_png(const std::string& path, int w, int h, const std::vector<uint8_t>& rgb, std::string* err) {
    if (rgb.size() != size_t(w) * size_t(h) * 3) {
        if (err) *err = "bad buffer size";
        return false;
    }
    std::vector<uint8_t> raw;
    raw.reserve(size_t(h) * (size_t(w) * 3 + 1));
    for (int y = 0; y < h; ++y) {
        raw.push_back(0);  // filter: none
        const auto row = rgb.begin() + std::ptrdiff_t(size_t(y) * size_t(w) * 3);
        raw.insert(raw.end(), row, row + std::ptrdiff_t(size_t(w) * 3));
    }
    std::vector<uint8_t> z = {0x78, 0x01};
    for (size_t pos = 0; pos < raw.size() || raw.empty();) {
        const size_t n = std::min<size_t>(65535, raw.size() - pos);
        const bool last = pos + n == raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back(uint8_t(n & 0xFF));
        z.push_back(uint8_t(n >> 8));
        z.push_back(uint8_t(~n & 0xFF));
        z.push_back(uint8_t((~n >> 8) & 0xFF));
        z.insert(z.end(), raw.begin() + std::ptrdiff_t(pos), raw.begin() + std::ptrdiff_t(pos + n));
        pos += n;
        if (last) break;
    }
    uint32_t a = 1, b = 0;
    for (uint8_t v : raw) {
        a = (a + v) % 65521;
        b = (b + a) % 65521;
    }
    be32(z, (b << 16) | a);
    std::vector<uint8_t> png = {137, 80, 78, 71, 13, 10, 26, 10};
    std::vector<uint8_t> ihdr;
    be32(ihdr, uint32_t(w));
    be32(ihdr, uint32_t(h));
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});
    chunk(png, "IHDR", ihdr);
    chunk(png, "IDAT", z);
    chunk(png, "IEND", {});
    std::ofstream f(path, std::ios::binary);
    if (!f) {
        if (err) *err = "cannot open " + path;
        return false;
    }
    f.write(reinterpret_cast<const char*>(png.data()), std::streamsize(png.size()));
    return bool(f);
}

bool have_ffmpeg() { return std::system("ffmpeg -hide_banner -loglevel error -version >/dev/null 2>&1") == 0; }

bool encode_mp4(const std::string& frames_dir, const std::string& basename, const std::string& out_mp4, int fps,
                std::string* err) {
    if (!have_ffmpeg()) {
        if (err) *err = "ffmpeg not found in PATH";
        return false;
    }
    std::ostringstream cmd;
    cmd << "ffmpeg -hide_banner -loglevel error -y -framerate " << fps << " -i \""
        << (fs::path(frames_dir) / (basename + "%04d.ppm")).string()
        << "\" -vf \"scale=trunc(iw/2)*2:trunc(ih/2)*2\" -c:v libx264 -pix_fmt yuv420p -movflags +faststart \""
        << out_mp4 << "\"";
    if (std::system(cmd.str().c_str()) != 0) {
        if (err) *err = "ffmpeg failed";
        return false;
    }
    return true;
}

}  // namespace raymotion
