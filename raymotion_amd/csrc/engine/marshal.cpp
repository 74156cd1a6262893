// Scene marshalling and render-switch handling in front of the C-ABI.
#include "marshal.hpp"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <new>
#include <cstdlib>
#include <string>
#include <unordered_map>

#include "rm_internal.h"
#include "rm_layout.h"
#include "rm_parallel.h"

namespace raymotion {
namespace {

rm_vec3 pack(const Vec3& v) { return {float(v.x), float(v.y), float(v.z)}; }

bool env_is(const char* name, const char* value) {
    const char* v = std::getenv(name);
    return v && std::string(v) == value;
}

}  // namespace

rm_scene_desc MarshalledScene::desc(const Scene& scene) const {
    rm_scene_desc d{};
    d.tris = tris.data();
    d.ntri = tris.size();
    d.nodes = nodes.data();
    d.nnode = nodes.size();
    d.bvh_tri = scene.bvh_tri.data();
    d.light_tri = scene.light_tri.data();
    d.light_cdf = cdf.data();
    d.nlight = scene.light_tri.size();
    d.light_total = float(scene.light_total);
    d.texels = texels.data();
    d.ntexel = texels.size() / 3;
    return d;
}

rmkeys::RecordKeys MarshalledScene::keys() const {
    rmkeys::RecordKeys k;
    k.kp = kp.data();
    k.kn = kn.data();
    k.km = km.data();
    k.chunks = kp.size();
    return k;
}

bool marshal_scene(const Scene& scene, MarshalledScene& out, GeometrySink* sink, bool with_tree) {
    const size_t n = scene.tris.size(), old_n = out.tris.size();
    const size_t kept = std::min(old_n, n);  // records of the last marshal still in place
    out.tris.resize(n);
    out.texels.clear();
    if (sink) {
        out.geo.resize(n * rmh::kTriGeoWords);
        out.nrm.resize(n * rmh::kTriNrmWords);
        out.nn.resize(n);
    }
    std::atomic<bool> changed{old_n != n}, sink_ok{true};
    // Field conversion in parallel chunks (the records' key chunks). A record
    // equal to the one already in place is not written again, so re-marshalling
    // an unchanged scene only reads; each chunk's keys are taken from the
    // records as they are converted. Each chunk lists its textured triangles,
    // which then get atlas offsets in triangle order (their chunks' keys are
    // taken after that).
    constexpr size_t kChunk = rmkeys::kKeyChunk;
    const size_t chunks = rmkeys::key_chunks(n);
    out.kp.resize(chunks);
    out.kn.resize(chunks);
    out.km.resize(chunks);
    std::vector<std::vector<uint32_t>> textured(chunks);
    auto store = [&](size_t i, const rm_triangle& t) {
        rm_triangle& d = out.tris[i];
        if (i >= kept || std::memcmp(&d, &t, sizeof t) != 0) {
            d = t;
            return true;
        }
        return false;
    };
    rmpar::parallel_chunks(chunks, [&](size_t c) {
        const size_t a = c * kChunk, b = std::min(n, a + kChunk);
        rmkeys::RecordMix mix(c);
        bool dirty = false;
        for (size_t i = a; i < b; ++i) {
            const Triangle& s = scene.tris[i];
            rm_triangle t;
            t.v0 = pack(s.v0); t.v1 = pack(s.v1); t.v2 = pack(s.v2);
            t.n = pack(s.n); t.nn = pack(s.nn);
            t.n0 = pack(s.n0); t.n1 = pack(s.n1); t.n2 = pack(s.n2);
            t.ar = float(s.ar); t.ag = float(s.ag); t.ab = float(s.ab);
            t.er = float(s.er); t.eg = float(s.eg); t.eb = float(s.eb);
            t.ior = float(s.ior); t.rough = float(s.rough); t.metallic = float(s.metallic);
            t.alpha = float(s.alpha);
            t.shader = s.shader;
            t.tu0 = float(s.tu0); t.tv0 = float(s.tv0); t.tu1 = float(s.tu1);
            t.tv1 = float(s.tv1); t.tu2 = float(s.tu2); t.tv2 = float(s.tv2);
            t.texture_offset = -1;
            t.texture_width = 0;
            t.texture_height = 0;
            if (s.texture) {
                // atlas fields as in place, until the atlas pass below
                if (i < kept) {
                    const rm_triangle& d = out.tris[i];
                    t.texture_offset = d.texture_offset;
                    t.texture_width = d.texture_width;
                    t.texture_height = d.texture_height;
                }
                textured[c].push_back(uint32_t(i));
            }
            dirty |= store(i, t);
            mix.add(reinterpret_cast<const unsigned char*>(&t));
            if (sink) {
                rmh::pack_tri_geo(t, int32_t(i), &out.geo[rmh::kTriGeoWords * i]);
                rmh::pack_tri_nrm(t, &out.nrm[rmh::kTriNrmWords * i], &out.nn[i]);
            }
        }
        out.kp[c] = mix.p.done();
        out.kn[c] = mix.n.done();
        out.km[c] = mix.m.done();
        if (dirty) changed.store(true);
        if (sink && !sink->put(a, b - a, &out.geo[rmh::kTriGeoWords * a], &out.nrm[rmh::kTriNrmWords * a],
                               &out.nn[a]))
            sink_ok.store(false);
    });
    std::unordered_map<const Texture*, int> atlas;
    for (const auto& list : textured)
        for (uint32_t i : list) {
            const Triangle& s = scene.tris[i];
            auto ins = atlas.emplace(s.texture.get(), int(out.texels.size() / 3));
            if (ins.second)
                for (const Vec3& p : s.texture->pixels) {
                    out.texels.push_back(float(p.x));
                    out.texels.push_back(float(p.y));
                    out.texels.push_back(float(p.z));
                }
            rm_triangle t = out.tris[i];
            t.texture_offset = ins.first->second;
            t.texture_width = s.texture->width;
            t.texture_height = s.texture->height;
            if (store(i, t)) changed.store(true);
        }
    if (!atlas.empty()) rmpar::parallel_chunks(chunks, [&](size_t c) {
        if (textured[c].empty()) return;
        const size_t a = c * kChunk, b = std::min(n, a + kChunk);
        rmkeys::RecordMix mix(c);
        for (size_t i = a; i < b; ++i) mix.add(reinterpret_cast<const unsigned char*>(&out.tris[i]));
        out.kp[c] = mix.p.done();
        out.kn[c] = mix.n.done();
        out.km[c] = mix.m.done();
    });
    // Preorder with escape links: a node is followed by its left subtree, then
    // its right subtree; `right` holds the index just past its own subtree.
    // Not needed for an upload that builds its own tree.
    out.nodes.clear();
    if (with_tree) out.nodes.reserve(scene.nodes.size());
    struct Item { int index; bool exit; };
    std::vector<Item> todo;
    if (with_tree && !scene.nodes.empty()) todo.push_back({0, false});
    while (!todo.empty()) {
        const Item it = todo.back();
        todo.pop_back();
        if (it.exit) {
            out.nodes[size_t(it.index)].right = int32_t(out.nodes.size());
            continue;
        }
        const BvhNode& s = scene.nodes.at(size_t(it.index));
        rm_bvh_node n{};
        n.mn = pack(s.mn);
        n.mx = pack(s.mx);
        n.offset = s.offset;
        n.count = s.count;
        const int at = int(out.nodes.size());
        out.nodes.push_back(n);
        todo.push_back({at, true});
        if (!s.is_leaf()) {
            todo.push_back({s.right, false});
            todo.push_back({s.left, false});
        }
    }
    out.cdf.assign(scene.light_cdf.begin(), scene.light_cdf.end());
    out.changed = changed.load();
    return sink_ok.load();
}

namespace {

struct StageSink final : GeometrySink {
    rm_ctx* ctx;
    explicit StageSink(rm_ctx* c) : ctx(c) {}
    bool put(size_t first, size_t count, const rmh::F4* geo, const rmh::F4* nrm, const rmh::F4* nn) override {
        return rm_geo_stage_put(ctx, first, count, geo, nrm, nn);
    }
};

}  // namespace

int marshal_and_upload(rm_ctx* ctx, const Scene& scene, MarshalledScene& m, bool gpu_tree, int rebuild) {
    if (!ctx) return RM_ERR_INVALID;
    bool staged = false;
    try {
        if (gpu_tree && m.changed && !scene.tris.empty() && rm_geo_stage_begin(ctx, scene.tris.size()) == RM_OK) {
            StageSink sink(ctx);
            staged = marshal_scene(scene, m, &sink, false);
        } else {
            marshal_scene(scene, m, nullptr, !gpu_tree);
        }
    } catch (const std::bad_alloc&) {
        // Host or pinned memory for the records. `m` stays allocated (staging
        // copies from it may be queued); the next marshal rewrites it.
        return RM_ERR_ALLOC;
    }
    const rm_scene_desc d = m.desc(scene);
    return gpu_tree ? rm_set_scene_gpu_tree_keyed(ctx, &d, rebuild, m.keys(), staged)
                    : rm_set_scene_keyed(ctx, &d, m.keys());
}

rm_frame make_frame(const Camera& cam, const RenderConfig& cfg) {
    rm_frame f{};
    f.camera = {float(cam.x), float(cam.y), float(cam.z), float(cam.focal), float(cam.m0), float(cam.m1),
                float(cam.m2), float(cam.m3), float(cam.m4), float(cam.m5), float(cam.m6), float(cam.m7),
                float(cam.m8)};
    f.width = cfg.width;
    f.height = cfg.height;
    f.spp = cfg.spp;
    f.bounces = cfg.max_bounces;
    f.resolution = float(cfg.resolution);
    f.clamp_value = float(cfg.clamp);
    f.adapt_rel = float(cfg.adapt_rel);
    f.adapt_abs = float(cfg.adapt_abs);
    f.adapt_min = std::getenv("UOW2_NOADAPT") ? 0 : cfg.adapt_min;
    f.adapt_step = cfg.adapt_step > 0 ? cfg.adapt_step : 8;
    f.nomis = std::getenv("UOW2_NOMIS") != nullptr;
    const bool restir_off = env_is("RAYMOTION_RESTIR", "0");
    f.restir_candidates = restir_off ? 1 : (cfg.restir_candidates > 0 ? cfg.restir_candidates : 1);
    f.restir_candidates = std::min(32, std::max(1, f.restir_candidates));
    f.restir_spatial = restir_off ? 0 : (cfg.restir_spatial > 0 ? cfg.restir_spatial : 0);
    f.restir_spatial = std::min(8, std::max(0, f.restir_spatial));
    f.restir_mcap = cfg.restir_mcap > 0 ? cfg.restir_mcap : 128;
    f.restir_radius = cfg.restir_radius > 0 ? float(cfg.restir_radius) : 16.0f;
    if (const char* v = std::getenv("RAYMOTION_RESTIR_CANDIDATES")) f.restir_candidates = std::max(1, std::atoi(v));
    if (const char* v = std::getenv("RAYMOTION_RESTIR_SPATIAL")) f.restir_spatial = std::max(0, std::atoi(v));
    f.denoise = env_is("RAYMOTION_DENOISE", "0") ? 0 : 1;
    f.denoise_iterations = std::max(0, cfg.denoise_iterations);
    f.temporal = env_is("RAYMOTION_RESTIR_TEMPORAL", "0") ? 0 : 1;
    return f;
}

}  // namespace raymotion
