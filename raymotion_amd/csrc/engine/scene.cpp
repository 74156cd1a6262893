// Triangle precompute, light CDF and texture lookup.
// Behaviour follows src/engine/scene.cpp:9-87 of the reference: the same FP64
// expressions in the same evaluation order, so bounds, normals and the CDF the
// GPU receives are bit-identical to what the reference engine produces.
#include "raymotion/scene.hpp"

#include <algorithm>
#include <cmath>

namespace raymotion {

Vec3 Texture::sample(double u, double v) const {
    u -= std::floor(u);
    v -= std::floor(v);
    const int x = std::min(int(u * width), width - 1);
    const int y = height - 1 - std::min(int(v * height), height - 1);
    return pixels[size_t(y) * size_t(width) + size_t(x)];
}

bool Scene::precompute_triangle(const Vec3& p0, const Vec3& p1, const Vec3& p2, Triangle& t, double& gram) {
    const Vec3 e1 = p1 - p0;
    const Vec3 e2 = p2 - p0;
    const double g00 = dot(e1, e1), g01 = dot(e1, e2), g11 = dot(e2, e2);
    gram = g00 * g11 - g01 * g01;  // |e1 x e2|^2
    if (std::abs(gram) <= 1e-12) return false;
    t.v0 = p0; t.v1 = p1; t.v2 = p2;
    t.n = cross(e1, e2);
    t.d = -(t.n.x * p0.x + (t.n.y * p0.y + t.n.z * p0.z));
    const double inv = 1.0 / gram;
    t.U = {(e1.x * g11 - e2.x * g01) * inv, (e1.y * g11 - e2.y * g01) * inv,
           (e1.z * g11 - e2.z * g01) * inv};
    t.V = {(e2.x * g00 - e1.x * g01) * inv, (e2.y * g00 - e1.y * g01) * inv,
           (e2.z * g00 - e1.z * g01) * inv};
    t.denom = gram;
    t.nn = normalize(t.n);
    t.mn = {std::min({p0.x, p1.x, p2.x}), std::min({p0.y, p1.y, p2.y}), std::min({p0.z, p1.z, p2.z})};
    t.mx = {std::max({p0.x, p1.x, p2.x}), std::max({p0.y, p1.y, p2.y}), std::max({p0.z, p1.z, p2.z})};
    return true;
}

namespace {
// Light bookkeeping of one triangle (lights are picked proportionally to
// area x luminance).
void add_light(Scene& s, int index, double er, double eg, double eb, double gram) {
    if (er + (eg + eb) > 0) {
        const double area = 0.5 * std::sqrt(std::max(gram, 0.0));
        if (area > 1e-12) {
            double lum = 0.2126 * er + 0.7152 * eg + 0.0722 * eb;
            if (lum < 1e-6) lum = 1e-6;
            s.light_total += area * lum;
            s.light_tri.push_back(index);
            s.light_cdf.push_back(s.light_total);
        }
    }
}
}  // namespace

int Scene::add_triangle(const Vec3& p0, const Vec3& p1, const Vec3& p2, double u0, double v0,
                        double u1, double v1, double u2, double v2, const Vec3& n0,
                        const Vec3& n1, const Vec3& n2, double ar, double ag, double ab,
                        double er, double eg, double eb, double metallic, double ior,
                        double rough, int shader, double alpha) {
    Triangle t;
    double gram = 0;
    if (!precompute_triangle(p0, p1, p2, t, gram)) return -1;
    t.tu0 = u0; t.tv0 = v0; t.tu1 = u1; t.tv1 = v1; t.tu2 = u2; t.tv2 = v2;
    t.n0 = n0; t.n1 = n1; t.n2 = n2;
    t.ar = ar; t.ag = ag; t.ab = ab;
    t.er = er; t.eg = eg; t.eb = eb;
    t.ior = ior;
    t.alpha = std::isfinite(alpha) ? std::clamp(alpha, 0.0, 1.0) : 1.0;
    t.rough = rough;
    t.shader = shader;
    t.metallic = metallic;
    const int index = int(tris.size());
    tris.push_back(std::move(t));
    add_light(*this, index, er, eg, eb, gram);
    return index;
}

void Scene::rebuild_lights() {
    light_tri.clear();
    light_cdf.clear();
    light_total = 0;
    for (size_t i = 0; i < tris.size(); ++i) {
        const Triangle& t = tris[i];
        add_light(*this, int(i), t.er, t.eg, t.eb, t.denom);
    }
}

}  // namespace raymotion
