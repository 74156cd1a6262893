// OBJ/MTL loader for scene programs (input side of the engine; the GPU path
// only sees the resulting triangles). Conventions follow src/engine/obj.cpp of
// the reference: z is negated on input (right- to left-handed), faces are fanned
// as (0, i+1, i), missing normals fall back to the geometric normal, and a
// texture is attached only when all three corners carry UVs.
#include "raymotion/obj.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <unordered_map>

#include "png.h"

namespace raymotion {
namespace {

struct Material {
    Vec3 kd{1, 1, 1}, ke{};
    double ior = 1, rough = 0.5, metallic = 0;
    std::shared_ptr<const Texture> texture;
};
using Materials = std::unordered_map<std::string, Material>;

std::string read_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::ostringstream s;
    s << f.rdbuf();
    return s.str();
}

double srgb_to_linear(double c) { return c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4); }

// Binary (P6) or ASCII (P3) PPM, 8-bit.
std::shared_ptr<Texture> load_ppm(const std::string& data, const std::string& path) {
    std::istringstream in(data);
    std::string magic;
    in >> magic;
    if (magic != "P6" && magic != "P3") throw std::runtime_error("unsupported texture format (PNG or PPM): " + path);
    auto skip = [&] {
        in >> std::ws;
        while (in.peek() == '#') {
            std::string line;
            std::getline(in, line);
            in >> std::ws;
        }
    };
    int w = 0, h = 0, maxv = 0;
    skip(); in >> w;
    skip(); in >> h;
    skip(); in >> maxv;
    if (!in || w <= 0 || h <= 0 || maxv <= 0 || maxv > 255) throw std::runtime_error("invalid PPM header: " + path);
    auto tex = std::make_shared<Texture>();
    tex->width = w;
    tex->height = h;
    tex->pixels.resize(size_t(w) * size_t(h));
    if (magic == "P6") {
        in.get();
        std::vector<unsigned char> raw(size_t(w) * size_t(h) * 3);
        in.read(reinterpret_cast<char*>(raw.data()), std::streamsize(raw.size()));
        if (size_t(in.gcount()) != raw.size()) throw std::runtime_error("truncated PPM: " + path);
        for (size_t i = 0; i < tex->pixels.size(); ++i)
            tex->pixels[i] = {srgb_to_linear(raw[3 * i] / double(maxv)), srgb_to_linear(raw[3 * i + 1] / double(maxv)),
                              srgb_to_linear(raw[3 * i + 2] / double(maxv))};
    } else {
        for (auto& p : tex->pixels) {
            int r, g, b;
            if (!(in >> r >> g >> b)) throw std::runtime_error("truncated PPM: " + path);
            p = {srgb_to_linear(r / double(maxv)), srgb_to_linear(g / double(maxv)), srgb_to_linear(b / double(maxv))};
        }
    }
    return tex;
}

// A map_Kd image: PNG (any colour type, reduced to 8-bit RGB as stb_image
// does for the reference) or PPM; texels are the sRGB-decoded bytes.
std::shared_ptr<Texture> load_texture(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot load texture: " + path);
    std::ostringstream s;
    s << f.rdbuf();
    const std::string data = s.str();
    if (!is_png(data)) return load_ppm(data, path);
    Rgb8Image img;
    try {
        img = decode_png(data);
    } catch (const std::exception& e) {
        throw std::runtime_error("cannot load texture: " + path + " (" + e.what() + ")");
    }
    auto tex = std::make_shared<Texture>();
    tex->width = img.width;
    tex->height = img.height;
    tex->pixels.resize(size_t(img.width) * size_t(img.height));
    for (size_t i = 0; i < tex->pixels.size(); ++i)
        tex->pixels[i] = {srgb_to_linear(img.rgb[3 * i] / 255.0), srgb_to_linear(img.rgb[3 * i + 1] / 255.0),
                          srgb_to_linear(img.rgb[3 * i + 2] / 255.0)};
    return tex;
}

Materials parse_mtl(const std::string& path) {
    Materials out;
    std::unordered_map<std::string, std::shared_ptr<const Texture>> textures;
    std::istringstream in(read_file(path));
    const std::filesystem::path dir = std::filesystem::path(path).parent_path();
    std::string line;
    Material* cur = nullptr;
    while (std::getline(in, line)) {
        line = line.substr(0, line.find('#'));
        std::istringstream row(line);
        std::string key;
        if (!(row >> key)) continue;
        if (key == "newmtl") {
            std::string name;
            std::getline(row >> std::ws, name);
            while (!name.empty() && (name.back() == '\r' || name.back() == ' ')) name.pop_back();
            cur = &out[name];
        } else if (!cur) {
            continue;
        } else if (key == "Kd") {
            row >> cur->kd.x >> cur->kd.y >> cur->kd.z;
        } else if (key == "Ke") {
            row >> cur->ke.x >> cur->ke.y >> cur->ke.z;
        } else if (key == "Ni") {
            row >> cur->ior;
        } else if (key == "Pr") {
            row >> cur->rough;
        } else if (key == "Pm") {
            row >> cur->metallic;
        } else if (key == "map_Kd") {
            // One path, optionally quoted, backslashes read as separators
            // (src/engine/obj.cpp:108-113); each file is decoded once per MTL.
            std::string file;
            std::getline(row >> std::ws, file);
            while (!file.empty() && (file.back() == '\r' || file.back() == ' ')) file.pop_back();
            if (file.empty() || file[0] == '-')
                throw std::runtime_error("map_Kd requires a texture path (options unsupported): " + path);
            if (file.size() > 1 && file.front() == '"' && file.back() == '"') file = file.substr(1, file.size() - 2);
            std::replace(file.begin(), file.end(), '\\', '/');
            const std::string full = (dir / file).lexically_normal().string();
            auto found = textures.find(full);
            if (found == textures.end()) found = textures.emplace(full, load_texture(full)).first;
            cur->texture = found->second;
        }
    }
    return out;
}

int parse_obj(Scene& scene, const std::string& data, double ox, double oy, double oz, double scale,
              const Material& fallback, int shader, const Materials& materials) {
    std::vector<Vec3> verts, norms, uvs;
    std::istringstream in(data);
    std::string line;
    Material mat = fallback;
    int added = 0;
    auto index = [](const std::string& text, size_t size) {
        size_t used = 0;
        const int i = std::stoi(text, &used);
        if (used != text.size() || i == 0) throw std::runtime_error("invalid OBJ index");
        const int r = i > 0 ? i - 1 : int(size) + i;
        if (r < 0 || r >= int(size)) throw std::runtime_error("OBJ index out of range");
        return r;
    };
    while (std::getline(in, line)) {
        line = line.substr(0, line.find('#'));
        while (!line.empty() && (line.back() == '\r' || line.back() == ' ' || line.back() == '\t')) line.pop_back();
        std::istringstream row(line);
        std::string kind;
        if (!(row >> kind)) continue;
        if (kind == "v" || kind == "vn" || kind == "vt") {
            Vec3 v;
            if (!(row >> v.x >> v.y)) throw std::runtime_error("invalid OBJ vertex");
            if (kind != "vt" && !(row >> v.z)) throw std::runtime_error("invalid OBJ vertex");
            if (!std::isfinite(v.x) || !std::isfinite(v.y) || !std::isfinite(v.z))
                throw std::runtime_error("non-finite OBJ vertex");
            if (kind == "v") verts.push_back({ox + v.x * scale, oy + v.y * scale, oz - v.z * scale});
            else if (kind == "vn") norms.push_back({v.x, v.y, -v.z});
            else uvs.push_back(v);
        } else if (kind == "usemtl") {
            std::string name;
            std::getline(row >> std::ws, name);
            const auto it = materials.find(name);
            mat = it == materials.end() ? fallback : it->second;
        } else if (kind == "f") {
            struct Corner { int v = -1, t = -1, n = -1; };
            std::vector<Corner> face;
            std::string tok;
            while (row >> tok) {
                std::array<std::string, 3> parts;
                std::istringstream fields(tok);
                for (auto& p : parts) std::getline(fields, p, '/');
                Corner c;
                c.v = index(parts[0], verts.size());
                if (!parts[1].empty()) c.t = index(parts[1], uvs.size());
                if (!parts[2].empty()) c.n = index(parts[2], norms.size());
                face.push_back(c);
            }
            if (face.size() < 3) throw std::runtime_error("OBJ face needs three vertices");
            for (size_t i = 1; i + 1 < face.size(); ++i) {
                const Corner cs[3] = {face[0], face[i + 1], face[i]};
                Vec3 p[3], n[3], t[3];
                for (int j = 0; j < 3; ++j) {
                    p[j] = verts[size_t(cs[j].v)];
                    if (cs[j].t >= 0) t[j] = uvs[size_t(cs[j].t)];
                }
                const Vec3 g = normalize(cross(p[1] - p[0], p[2] - p[0]));
                for (int j = 0; j < 3; ++j) n[j] = cs[j].n >= 0 ? norms[size_t(cs[j].n)] : g;
                const int id = scene.add_triangle(p[0], p[1], p[2], t[0].x, t[0].y, t[1].x, t[1].y, t[2].x, t[2].y,
                                                  n[0], n[1], n[2], mat.kd.x, mat.kd.y, mat.kd.z, mat.ke.x, mat.ke.y,
                                                  mat.ke.z, mat.metallic, mat.ior, mat.rough, shader);
                if (id >= 0) {
                    if (cs[0].t >= 0 && cs[1].t >= 0 && cs[2].t >= 0) scene.tris[size_t(id)].texture = mat.texture;
                    ++added;
                }
            }
        }
    }
    return added;
}

}  // namespace

int add_obj_tokens(Scene& scene, const std::string& data, double ox, double oy, double oz, double scale, double ar,
                   double ag, double ab, double er, double eg, double eb, double ior, double rough, int shader,
                   double metallic) {
    Material m;
    m.kd = {ar, ag, ab};
    m.ke = {er, eg, eb};
    m.ior = ior;
    m.rough = rough;
    m.metallic = metallic;
    return parse_obj(scene, data, ox, oy, oz, scale, m, shader, {});
}

int add_obj_file(Scene& scene, const std::string& path, double ox, double oy, double oz, double scale, double ar,
                 double ag, double ab, double er, double eg, double eb, double ior, double rough, int shader,
                 double metallic) {
    std::string data;
    try {
        data = read_file(path);
    } catch (const std::runtime_error&) {
        return -1;
    }
    return add_obj_tokens(scene, data, ox, oy, oz, scale, ar, ag, ab, er, eg, eb, ior, rough, shader, metallic);
}

int load_obj_materials(Scene& scene, const std::string& obj, const std::string& mtl) {
    const Materials mats = parse_mtl(mtl);
    Material fallback;
    return parse_obj(scene, read_file(obj), 0, 0, 0, 1, fallback, 0, mats);
}

}  // namespace raymotion
