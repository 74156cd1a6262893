// Scene packing: reference-layout arrays -> device layouts (see rm_layout.h),
// and the content keys that decide what a new scene re-uploads.
#include "rm_pack.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>

#include "rm_layout.h"
#include "rm_parallel.h"

namespace rmh {
namespace {

inline float ibf(int32_t v) {
    float f;
    std::memcpy(&f, &v, 4);
    return f;
}

struct BinTree {  // binary tree recovered from the preorder array
    const rm_bvh_node* nodes;
    size_t n;
    int left(int i) const { return i + 1; }
    int right(int i) const { return nodes[i + 1].right; }
    bool leaf(int i) const { return nodes[i].count > 0; }
    float area(int i) const {
        const float dx = nodes[i].mx.x - nodes[i].mn.x, dy = nodes[i].mx.y - nodes[i].mn.y,
                    dz = nodes[i].mx.z - nodes[i].mn.z;
        return dx * dy + dy * dz + dz * dx;
    }
};

// One 4-wide node during collapse: its children's binary nodes and slots,
// and the next slot whose inner child still has to be collapsed.
struct CollapseFrame {
    int dst, depth, stack_below, next;
    int nk;
    int kids[kWide];
    float lo[3][kWide], hi[3][kWide];
    int ref[kWide], kind[kWide];
};

// Gathers the children of wide node `dst` from the binary subtree under `b`:
// repeatedly opens the inner child with the largest surface area until all
// kWide slots are used (a leaf root becomes the single child of the root).
void gather(const BinTree& t, int b, int dst, int depth, int stack_here, CollapseFrame& f, PackedScene& out) {
    f.dst = dst;
    f.depth = depth;
    f.next = 0;
    f.nk = 0;
    if (t.leaf(b)) {
        f.kids[f.nk++] = b;
    } else {
        f.kids[f.nk++] = t.left(b);
        f.kids[f.nk++] = t.right(b);
        while (f.nk < kWide) {
            int best = -1;
            float best_area = -1.0f;
            for (int k = 0; k < f.nk; ++k)
                if (!t.leaf(f.kids[k]) && t.area(f.kids[k]) > best_area) {
                    best_area = t.area(f.kids[k]);
                    best = k;
                }
            if (best < 0) break;
            const int open = f.kids[best];
            f.kids[best] = t.left(open);
            f.kids[f.nk++] = t.right(open);
        }
    }
    const float inf = std::numeric_limits<float>::infinity();
    int inner = 0;
    for (int c = 0; c < kWide; ++c) {
        if (c >= f.nk) {
            for (int a = 0; a < 3; ++a) { f.lo[a][c] = inf; f.hi[a][c] = -inf; }
            f.ref[c] = 0;
            f.kind[c] = -1;
            continue;
        }
        const rm_bvh_node& n = t.nodes[f.kids[c]];
        f.lo[0][c] = n.mn.x; f.lo[1][c] = n.mn.y; f.lo[2][c] = n.mn.z;
        f.hi[0][c] = n.mx.x; f.hi[1][c] = n.mx.y; f.hi[2][c] = n.mx.z;
        if (n.count > 0) {
            f.ref[c] = n.offset;
            f.kind[c] = n.count;
        } else {
            ++inner;
            f.ref[c] = -1;  // the child's wide node, assigned when it is allocated
            f.kind[c] = 0;
        }
    }
    // Deepest stack below: every inner child but the first visited is stacked.
    f.stack_below = stack_here + std::max(0, inner - 1);
    out.stack_need = std::max(out.stack_need, f.stack_below);
}

// Collapses the binary tree into 4-wide nodes, numbered in depth-first
// preorder (a node's first inner child follows it, each later one follows the
// previous one's subtree). Iterative, with one frame per level of the wide
// tree on the heap, so any depth of the caller's tree is handled.
void collapse(const BinTree& t, PackedScene& out) {
    std::vector<CollapseFrame> path(1);
    gather(t, 0, 0, 0, 0, path[0], out);
    while (!path.empty()) {
        CollapseFrame& f = path.back();
        while (f.next < f.nk && f.kind[f.next] != 0) ++f.next;
        if (f.next == f.nk) {
            write_node(&out.nodes[size_t(f.dst) * kNodeWords], f.lo, f.hi, f.ref, f.kind);
            path.pop_back();
            continue;
        }
        const int c = f.next++;
        const int child = int(out.nodes.size() / kNodeWords);
        out.nodes.resize(out.nodes.size() + kNodeWords);
        out.depth.push_back(f.depth + 1);
        f.ref[c] = child;
        const int b = f.kids[c], depth = f.depth + 1, stack_here = f.stack_below;
        path.emplace_back();  // may move the frames: f is not used after this
        gather(t, b, child, depth, stack_here, path.back(), out);
    }
}

using rmpar::parallel_chunks;

constexpr size_t kChunk = rmkeys::kKeyChunk;  // triangles per parallel chunk
using rmkeys::Mix4;

// Hash of a byte array: 1 MiB pieces hashed in parallel, folded in order.
uint64_t hash_span(const void* data, size_t n, uint64_t seed) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    constexpr size_t kPiece = size_t(1) << 20;
    const size_t pieces = (n + kPiece - 1) / kPiece;
    std::vector<uint64_t> part(pieces);
    parallel_chunks(pieces, [&](size_t c) {
        const size_t a = c * kPiece, len = std::min(kPiece, n - a);
        Mix4 m(seed ^ c);
        m.words(p + a, len / 8);
        uint64_t tail = 0;
        std::memcpy(&tail, p + a + (len & ~size_t(7)), len & 7);
        part[c] = Mix4::fold(m.done(), tail ^ len);
    });
    uint64_t h = seed ^ n;
    for (uint64_t v : part) h = Mix4::fold(h, v);
    return h;
}

}  // namespace

SceneKeys scene_keys_from_chunks(const rm_scene_desc& s, const rmkeys::RecordKeys& rk) {
    SceneKeys k;
    k.ntri = s.ntri;
    k.nnode = s.nnode;
    k.nlight = s.nlight;
    k.ntexel = s.ntexel;
    uint64_t hp = 0x51ED0001 ^ s.ntri, hn = 0x51ED0002 ^ s.ntri, hm = 0x51ED0003 ^ s.ntri;
    for (size_t c = 0; c < rk.chunks; ++c) {
        hp = Mix4::fold(hp, rk.kp[c]);
        hn = Mix4::fold(hn, rk.kn[c]);
        hm = Mix4::fold(hm, rk.km[c]);
    }
    k.positions = hp;
    k.normals = hn;
    k.material = hash_span(s.texels, s.ntexel * 3 * sizeof(float), hm ^ s.ntexel);
    k.tree = hash_span(s.bvh_tri, (s.nnode > 0 ? s.ntri : 0) * sizeof(int32_t),
                       hash_span(s.nodes, s.nnode * sizeof(rm_bvh_node), 0x7EE ^ s.nnode));
    uint32_t lt;
    std::memcpy(&lt, &s.light_total, 4);
    k.lights = hash_span(s.light_cdf, s.nlight * sizeof(float),
                         hash_span(s.light_tri, s.nlight * sizeof(int32_t), 0x11647 ^ s.nlight) ^ lt);
    return k;
}

SceneKeys scene_keys(const rm_scene_desc& s) {
    static_assert(sizeof(rm_triangle) == 176, "triangle record");
    const size_t chunks = rmkeys::key_chunks(s.ntri);
    std::vector<uint64_t> kp(chunks), kn(chunks), km(chunks);
    const unsigned char* base = reinterpret_cast<const unsigned char*>(s.tris);
    parallel_chunks(chunks, [&](size_t c) {
        rmkeys::RecordMix mix(c);
        const size_t a = c * kChunk, b = std::min(s.ntri, a + kChunk);
        for (size_t i = a; i < b; ++i) mix.add(base + i * sizeof(rm_triangle));
        kp[c] = mix.p.done();
        kn[c] = mix.n.done();
        km[c] = mix.m.done();
    });
    rmkeys::RecordKeys rk;
    rk.kp = kp.data();
    rk.kn = kn.data();
    rk.km = km.data();
    rk.chunks = chunks;
    return scene_keys_from_chunks(s, rk);
}

uint64_t history_key(const SceneKeys& k, float light_total) {
    uint32_t lt;
    std::memcpy(&lt, &light_total, 4);
    uint64_t h = k.ntri * 0x9E3779B97F4A7C15ULL ^ k.nlight * 0xBF58476D1CE4E5B9ULL ^ lt;
    h = Mix4::fold(h, k.positions);
    h = Mix4::fold(h, k.normals);
    return Mix4::fold(h, k.material);
}

bool validate_scene(const rm_scene_desc& s, bool tree, bool materials, bool lights, std::string& err) {
    if (s.ntri > 0 && !s.tris) { err = "scene has triangles but no triangle array"; return false; }
    if (s.ntri >= (size_t(1) << 31) || s.nnode >= (size_t(1) << 31)) { err = "scene too large"; return false; }
    if (s.nnode > 0 && (!s.nodes || !s.bvh_tri)) { err = "scene has BVH nodes but no node or index array"; return false; }
    if (s.nlight > 0 && (!s.light_tri || !s.light_cdf)) { err = "scene has lights but no light arrays"; return false; }
    if (s.ntexel > 0 && !s.texels) { err = "scene has texels but no texel array"; return false; }
    if (tree) {
        // Preorder escape links and leaf ranges.
        for (size_t i = 0; i < s.nnode; ++i) {
            const rm_bvh_node& n = s.nodes[i];
            if (n.right <= int32_t(i) || size_t(n.right) > s.nnode) { err = "BVH escape link out of range"; return false; }
            if (n.count < 0) { err = "negative BVH leaf count"; return false; }
            if (n.count > 0) {
                if (n.offset < 0 || size_t(n.offset) + size_t(n.count) > s.ntri) { err = "BVH leaf range out of range"; return false; }
                if (n.right != int32_t(i) + 1) { err = "BVH leaf escape link is not the next node"; return false; }
            } else if (i + 1 >= s.nnode || s.nodes[i + 1].right >= n.right) {
                err = "BVH inner node without two children";
                return false;
            }
        }
        if (s.nnode > 0 && size_t(s.nodes[0].right) != s.nnode) { err = "BVH root does not span the node array"; return false; }
        for (size_t k = 0; k < (s.nnode > 0 ? s.ntri : 0); ++k)
            if (s.bvh_tri[k] < 0 || size_t(s.bvh_tri[k]) >= s.ntri) { err = "bvh_tri entry out of range"; return false; }
    }
    if (lights)
        for (size_t k = 0; k < s.nlight; ++k)
            if (s.light_tri[k] < 0 || size_t(s.light_tri[k]) >= s.ntri) { err = "light index out of range"; return false; }
    if (materials)
        for (size_t i = 0; i < s.ntri; ++i) {
            const rm_triangle& t = s.tris[i];
            if (t.texture_offset >= 0 &&
                (t.texture_width <= 0 || t.texture_height <= 0 ||
                 size_t(t.texture_offset) + size_t(t.texture_width) * size_t(t.texture_height) > s.ntexel)) {
                err = "texture range out of range";
                return false;
            }
        }
    return true;
}

void pack_tree(const rm_scene_desc& s, PackedScene& out) {
    out.nodes.clear();
    out.depth.clear();
    out.stack_need = 0;
    if (s.nnode == 0) return;
    BinTree bt{s.nodes, s.nnode};
    out.nodes.reserve(s.nnode / 2 * kNodeWords);
    out.nodes.resize(kNodeWords);
    out.depth.push_back(0);
    collapse(bt, out);
}

bool pack_geometry_sliced(const rm_scene_desc& s, PackedScene& out, size_t slices,
                          const std::function<bool(size_t, size_t)>& sink) {
    out.geo.resize(s.ntri * kTriGeoWords);
    out.nrm.resize(s.ntri * kTriNrmWords);
    out.nn.resize(s.ntri);
    constexpr size_t kTask = 8192;  // triangles per parallel task
    const size_t tasks = (s.ntri + kTask - 1) / kTask;
    slices = std::max<size_t>(1, std::min(slices, tasks));
    const size_t per_slice = tasks > 0 ? (tasks + slices - 1) / slices : 1;
    // Tasks are claimed in order; the worker that finishes a slice's last task
    // hands the slice to the sink (which must be callable from any thread).
    std::vector<std::atomic<size_t>> done((tasks + per_slice - 1) / per_slice);
    for (auto& d : done) d.store(0);
    std::atomic<bool> ok{true};
    parallel_chunks(tasks, [&](size_t t) {
        const size_t a = t * kTask, b = std::min(s.ntri, a + kTask);
        for (size_t k = a; k < b; ++k) {  // vertices in leaf order for traversal
            const int32_t j = s.nnode > 0 ? s.bvh_tri[k] : int32_t(k);
            pack_tri_geo(s.tris[j], j, &out.geo[3 * k]);
        }
        for (size_t i = a; i < b; ++i) pack_tri_nrm(s.tris[i], &out.nrm[3 * i], &out.nn[i]);  // scene order
        const size_t sl = t / per_slice, t0 = sl * per_slice, nt = std::min(per_slice, tasks - t0);
        if (done[sl].fetch_add(1) + 1 == nt && sink) {
            const size_t first = t0 * kTask, last = std::min(s.ntri, (t0 + nt) * kTask);
            if (!sink(first, last - first)) ok.store(false);
        }
    });
    return ok.load();
}

void pack_geometry(const rm_scene_desc& s, PackedScene& out) { (void)pack_geometry_sliced(s, out, 1, nullptr); }

void pack_lights(const rm_scene_desc& s, PackedScene& out) {
    out.lights.resize(s.nlight * kLightWords);
    out.cdf.assign(s.light_cdf, s.light_cdf + s.nlight);
    for (size_t k = 0; k < s.nlight; ++k) {  // CDF order
        const rm_triangle& t = s.tris[s.light_tri[k]];
        out.lights[4 * k] = {t.v0.x, t.v0.y, t.v0.z, t.v1.x};
        out.lights[4 * k + 1] = {t.v1.y, t.v1.z, t.v2.x, t.v2.y};
        out.lights[4 * k + 2] = {t.v2.z, t.nn.x, t.nn.y, t.nn.z};
        out.lights[4 * k + 3] = {t.er, t.eg, t.eb, t.alpha};
    }
    out.light_total = s.light_total;
}

void pack_materials(const rm_scene_desc& s, PackedScene& out) {
    constexpr int kKeyWords = 14;
    auto key_of = [](const rm_triangle& t, uint32_t* w) {
        const int32_t texoff = t.texture_offset >= 0 ? t.texture_offset : -1;
        const int32_t tw = texoff >= 0 ? t.texture_width : 0, th = texoff >= 0 ? t.texture_height : 0;
        const float v[kKeyWords] = {t.ar, t.ag, t.ab, t.alpha, t.er, t.eg, t.eb, t.ior, t.rough, t.metallic,
                                    ibf(t.shader), ibf(texoff), ibf(tw), ibf(th)};
        std::memcpy(w, v, sizeof(v));
    };
    // Per-triangle key hashes in parallel, then ids by first occurrence through
    // an open-addressing table (sequential, so ids do not depend on threads).
    const size_t n = s.ntri;
    std::vector<uint64_t> hk(n);
    const size_t chunks = (n + kChunk - 1) / kChunk;
    std::vector<uint8_t> chunk_opaque(chunks, 1), chunk_textured(chunks, 0);
    parallel_chunks(chunks, [&](size_t c) {
        const size_t a = c * kChunk, b = std::min(n, a + kChunk);
        for (size_t i = a; i < b; ++i) {
            uint32_t w[kKeyWords + 2] = {};
            key_of(s.tris[i], w);
            Mix4 m(0xA7);
            m.words(reinterpret_cast<const unsigned char*>(w), (kKeyWords + 1) / 2);
            hk[i] = m.done();
            if (!(s.tris[i].alpha >= 1.0f)) chunk_opaque[c] = 0;
            if (s.tris[i].texture_offset >= 0) chunk_textured[c] = 1;
        }
    });
    size_t cap = 1024;
    while (cap < 2 * n) cap <<= 1;
    std::vector<int32_t> slot(cap, -1);
    std::vector<uint32_t> keys;  // kKeyWords per material
    out.matid.resize(n);
    out.mats.clear();
    for (size_t i = 0; i < n; ++i) {
        uint32_t w[kKeyWords];
        key_of(s.tris[i], w);
        size_t at = size_t(hk[i]) & (cap - 1);
        int32_t id = -1;
        while (slot[at] >= 0) {
            if (std::memcmp(&keys[size_t(slot[at]) * kKeyWords], w, sizeof(w)) == 0) {
                id = slot[at];
                break;
            }
            at = (at + 1) & (cap - 1);
        }
        if (id < 0) {
            id = int32_t(keys.size() / kKeyWords);
            slot[at] = id;
            keys.insert(keys.end(), w, w + kKeyWords);
            float f[kKeyWords];
            std::memcpy(f, w, sizeof(w));
            out.mats.push_back({f[0], f[1], f[2], f[3]});
            out.mats.push_back({f[4], f[5], f[6], f[7]});
            out.mats.push_back({f[8], f[9], f[10], f[11]});
            out.mats.push_back({f[12], f[13], 0.0f, 0.0f});
        }
        out.matid[i] = id;
    }
    out.opaque = 1;
    bool textured = false;
    for (size_t c = 0; c < chunks; ++c) {
        if (!chunk_opaque[c]) out.opaque = 0;
        if (chunk_textured[c]) textured = true;
    }
    out.uv.clear();
    if (textured) {
        out.uv.resize(n * 2);
        for (size_t i = 0; i < n; ++i) {
            const rm_triangle& t = s.tris[i];
            out.uv[2 * i] = {t.tu0, t.tv0, t.tu1, t.tv1};
            out.uv[2 * i + 1] = {t.tu2, t.tv2, 0.0f, 0.0f};
        }
    }
    out.texels.assign(s.texels, s.texels + 3 * s.ntexel);
}

void topology_of(const rm_scene_desc& s, std::vector<int32_t>& bvh_tri, std::vector<int32_t>& topo) {
    bvh_tri.assign(s.bvh_tri, s.bvh_tri + (s.nnode > 0 ? s.ntri : 0));
    topo.resize(s.nnode * 3);
    for (size_t i = 0; i < s.nnode; ++i) {
        topo[3 * i] = s.nodes[i].right;
        topo[3 * i + 1] = s.nodes[i].offset;
        topo[3 * i + 2] = s.nodes[i].count;
    }
}

bool same_topology(const rm_scene_desc& s, const std::vector<int32_t>& prev_bvh_tri,
                   const std::vector<int32_t>& prev_topo) {
    if (prev_bvh_tri.size() != s.ntri || prev_topo.size() != s.nnode * 3 || s.nnode == 0) return false;
    if (std::memcmp(prev_bvh_tri.data(), s.bvh_tri, s.ntri * sizeof(int32_t)) != 0) return false;
    for (size_t i = 0; i < s.nnode; ++i)
        if (prev_topo[3 * i] != s.nodes[i].right || prev_topo[3 * i + 1] != s.nodes[i].offset ||
            prev_topo[3 * i + 2] != s.nodes[i].count)
            return false;
    return true;
}

}  // namespace rmh
