// The boundary's marshal takes the records' content keys while it writes them
// and skips records that did not change; the upload then uses those keys in
// place of hashing the records again. This program checks, over a scene
// re-marshalled after each kind of change (none, vertices, one material, a
// texture's pixels, more and fewer triangles), that the records equal a fresh
// marshal into empty arrays and that the keys equal the backend's own hash of
// the records (rmh::scene_keys), with and without the tree. Prints "ok".
#include <cstdio>
#include <cstring>
#include <memory>
#include <random>

#include "marshal.hpp"
#include "raymotion/bvh.hpp"
#include "rm_pack.h"

using namespace raymotion;

static int fails = 0;
#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) {                                                     \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);   \
            ++fails;                                                    \
        }                                                               \
    } while (0)

static void add_random(Scene& s, std::mt19937& g, int count, std::shared_ptr<const Texture> tex) {
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    for (int k = 0; k < count; ++k) {
        const Vec3 a{u(g) * 5, u(g) * 5, u(g) * 5};
        const Vec3 b{a.x + 0.5 + u(g) * 0.2, a.y, a.z + u(g) * 0.1}, c{a.x, a.y + 0.5, a.z + 0.3};
        const Vec3 nrm{0, 0, 1};
        const double e = k % 97 == 0 ? 3.0 : 0.0;
        const int i = s.add_triangle(a, b, c, 0, 0, 1, 0, 0, 1, nrm, nrm, nrm, 0.5 + 0.2 * u(g), 0.6, 0.7, e, e, e,
                                     0.3, 1.0, 0.4, k % 50 == 0 ? 1 : 0);
        if (i >= 0 && tex && k % 3 == 0) s.tris[size_t(i)].texture = tex;
    }
}

static void check(const Scene& s, MarshalledScene& m, const char* what) {
    marshal_scene(s, m);
    MarshalledScene fresh;
    marshal_scene(s, fresh);
    const bool same = m.tris.size() == fresh.tris.size() &&
                      std::memcmp(m.tris.data(), fresh.tris.data(), m.tris.size() * sizeof(rm_triangle)) == 0 &&
                      m.texels == fresh.texels;
    if (!same) std::printf("records differ after: %s\n", what);
    CHECK(same);
    rm_scene_desc d = m.desc(s);
    for (int tree = 0; tree < 2; ++tree) {
        if (!tree) {
            d.nodes = nullptr;
            d.nnode = 0;
            d.bvh_tri = nullptr;
        } else {
            d = m.desc(s);
        }
        const rmh::SceneKeys want = rmh::scene_keys(d);
        const rmh::SceneKeys got = rmh::scene_keys_from_chunks(d, m.keys());
        if (!(want == got)) std::printf("keys differ after: %s (tree %d)\n", what, tree);
        CHECK(want == got);
    }
}

int main() {
    std::mt19937 g(7);
    auto tex = std::make_shared<Texture>();
    tex->width = 4;
    tex->height = 3;
    tex->pixels.assign(12, Vec3{0.25, 0.5, 0.75});
    Scene s;
    add_random(s, g, 70000, tex);  // three key chunks, textured triangles in each
    build_bvh(s);
    MarshalledScene m;
    check(s, m, "first marshal");
    check(s, m, "no change");
    const rmh::SceneKeys k0 = rmh::scene_keys(m.desc(s));
    s.tris[40000].v1.x += 0.25;
    check(s, m, "one vertex");
    CHECK(!(rmh::scene_keys(m.desc(s)).positions == k0.positions));
    s.tris[65000].rough = 0.9;
    check(s, m, "one material");
    tex->pixels[5] = Vec3{1.0, 0.0, 0.0};
    check(s, m, "texture pixels");
    add_random(s, g, 500, tex);
    build_bvh(s);
    check(s, m, "more triangles");
    s.tris.resize(1000);
    s.light_tri.clear();
    s.light_cdf.clear();
    s.light_total = 0;
    build_bvh(s);
    check(s, m, "fewer triangles");
    s.tris.clear();
    s.clear_bvh();
    check(s, m, "empty");
    if (fails) return 1;
    std::printf("ok\n");
    return 0;
}
