// Scene packing: reference-layout arrays -> device layouts (see rm_layout.h).
#include "rm_pack.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <unordered_map>

#include "rm_layout.h"

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

// Collapses the binary subtree under `b` into 4-wide node `dst` (already
// allocated). Children are gathered by repeatedly opening the inner child with
// the largest surface area until four slots are used.
void collapse(const BinTree& t, int b, int dst, int depth, int stack_here, PackedScene& out) {
    int kids[4];
    int nk = 0;
    if (t.leaf(b)) {
        kids[nk++] = b;  // a leaf root becomes the single child of the root node
    } else {
        kids[nk++] = t.left(b);
        kids[nk++] = t.right(b);
        while (nk < 4) {
            int best = -1;
            float best_area = -1.0f;
            for (int k = 0; k < nk; ++k)
                if (!t.leaf(kids[k]) && t.area(kids[k]) > best_area) {
                    best_area = t.area(kids[k]);
                    best = k;
                }
            if (best < 0) break;
            const int open = kids[best];
            kids[best] = t.left(open);
            kids[nk++] = t.right(open);
        }
    }
    F4* w = &out.nodes[size_t(dst) * kNodeWords];
    const float inf = std::numeric_limits<float>::infinity();
    float lo[3][4], hi[3][4];
    int32_t ref[4], kind[4];
    int inner = 0;
    for (int c = 0; c < 4; ++c) {
        if (c >= nk) {
            for (int a = 0; a < 3; ++a) { lo[a][c] = inf; hi[a][c] = -inf; }
            ref[c] = 0;
            kind[c] = -1;
            continue;
        }
        const rm_bvh_node& n = t.nodes[kids[c]];
        lo[0][c] = n.mn.x; lo[1][c] = n.mn.y; lo[2][c] = n.mn.z;
        hi[0][c] = n.mx.x; hi[1][c] = n.mx.y; hi[2][c] = n.mx.z;
        if (n.count > 0) {
            ref[c] = n.offset;
            kind[c] = n.count;
        } else {
            ++inner;
            ref[c] = -1;  // patched below
            kind[c] = 0;
        }
    }
    // Deepest stack below: every inner child but the first visited is stacked.
    const int stack_below = stack_here + std::max(0, inner - 1);
    out.stack_need = std::max(out.stack_need, stack_below);
    for (int c = 0; c < nk; ++c) {
        if (kind[c] != 0) continue;
        const int child = int(out.nodes.size() / kNodeWords);
        out.nodes.resize(out.nodes.size() + kNodeWords);
        out.depth.push_back(depth + 1);
        ref[c] = child;
        collapse(t, kids[c], child, depth + 1, stack_below, out);
        w = &out.nodes[size_t(dst) * kNodeWords];  // storage may have moved
    }
    for (int a = 0; a < 3; ++a) {
        w[2 * a] = {lo[a][0], lo[a][1], lo[a][2], lo[a][3]};
        w[2 * a + 1] = {hi[a][0], hi[a][1], hi[a][2], hi[a][3]};
    }
    w[6] = {ibf(ref[0]), ibf(ref[1]), ibf(ref[2]), ibf(ref[3])};
    w[7] = {ibf(kind[0]), ibf(kind[1]), ibf(kind[2]), ibf(kind[3])};
}

struct MatKey {
    uint32_t w[14];
    bool operator==(const MatKey& o) const { return std::memcmp(w, o.w, sizeof(w)) == 0; }
};
struct MatHash {
    size_t operator()(const MatKey& k) const { return size_t(hash_bytes(k.w, sizeof(k.w), 0x51ED27)); }
};

}  // namespace

uint64_t hash_bytes(const void* data, size_t n, uint64_t h) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    size_t i = 0;
    for (; i + 8 <= n; i += 8) {
        uint64_t w;
        std::memcpy(&w, p + i, 8);
        h ^= w;
        h *= 0x100000001B3ULL;
        h ^= h >> 29;
    }
    for (; i < n; ++i) {
        h ^= p[i];
        h *= 0x100000001B3ULL;
    }
    return h ^ (h >> 32);
}

uint64_t history_key(const rm_scene_desc& s) {
    uint64_t h = uint64_t(s.ntri) * 0x9E3779B97F4A7C15ULL ^ uint64_t(s.nlight) * 0xBF58476D1CE4E5B9ULL;
    uint32_t lt;
    std::memcpy(&lt, &s.light_total, 4);
    h ^= lt;
    return hash_bytes(s.tris, s.ntri * sizeof(rm_triangle), h);
}

uint64_t content_key(const rm_scene_desc& s) {
    uint64_t h = history_key(s);
    h = hash_bytes(s.nodes, s.nnode * sizeof(rm_bvh_node), h ^ s.nnode);
    h = hash_bytes(s.bvh_tri, s.ntri * sizeof(int32_t), h);
    h = hash_bytes(s.light_tri, s.nlight * sizeof(int32_t), h);
    h = hash_bytes(s.light_cdf, s.nlight * sizeof(float), h);
    h = hash_bytes(s.texels, s.ntexel * 3 * sizeof(float), h ^ s.ntexel);
    return h;
}

void topology_of(const rm_scene_desc& s, std::vector<int32_t>& bvh_tri, std::vector<int32_t>& topo) {
    bvh_tri.assign(s.bvh_tri, s.bvh_tri + s.ntri);
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

bool pack_scene(const rm_scene_desc& s, PackedScene& out, std::string& err) {
    out = PackedScene{};
    if (s.ntri > 0 && !s.tris) { err = "scene has triangles but no triangle array"; return false; }
    if (s.ntri >= (size_t(1) << 31) || s.nnode >= (size_t(1) << 31)) { err = "scene too large"; return false; }
    if (s.nnode > 0 && (!s.nodes || !s.bvh_tri)) { err = "scene has BVH nodes but no node or index array"; return false; }
    if (s.nlight > 0 && (!s.light_tri || !s.light_cdf)) { err = "scene has lights but no light arrays"; return false; }
    // Tree validation: preorder escape links and leaf ranges.
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
    for (size_t k = 0; k < s.nlight; ++k)
        if (s.light_tri[k] < 0 || size_t(s.light_tri[k]) >= s.ntri) { err = "light index out of range"; return false; }
    for (size_t i = 0; i < s.ntri; ++i) {
        const rm_triangle& t = s.tris[i];
        if (t.texture_offset >= 0) {
            if (t.texture_width <= 0 || t.texture_height <= 0 ||
                size_t(t.texture_offset) + size_t(t.texture_width) * size_t(t.texture_height) > s.ntexel) {
                err = "texture range out of range";
                return false;
            }
        }
    }

    // 4-wide tree.
    if (s.nnode > 0) {
        BinTree bt{s.nodes, s.nnode};
        out.nodes.resize(kNodeWords);
        out.depth.push_back(0);
        collapse(bt, 0, 0, 0, 0, out);
    }

    // Triangles in leaf order for traversal (bvh_tri order).
    out.geo.resize(s.ntri * kTriGeoWords);
    for (size_t k = 0; k < s.ntri; ++k) {
        const int32_t j = s.nnode > 0 ? s.bvh_tri[k] : int32_t(k);
        const rm_triangle& t = s.tris[j];
        out.geo[3 * k] = {t.v0.x, t.v0.y, t.v0.z, t.v1.x};
        out.geo[3 * k + 1] = {t.v1.y, t.v1.z, t.v2.x, t.v2.y};
        out.geo[3 * k + 2] = {t.v2.z, ibf(j), 0.0f, 0.0f};
    }
    // Shading data in scene order.
    out.nrm.resize(s.ntri * kTriNrmWords);
    out.nn.resize(s.ntri);
    out.matid.resize(s.ntri);
    bool textured = false;
    std::unordered_map<MatKey, int32_t, MatHash> mat_index;
    for (size_t i = 0; i < s.ntri; ++i) {
        const rm_triangle& t = s.tris[i];
        out.nrm[3 * i] = {t.n0.x, t.n0.y, t.n0.z, t.n.x};
        out.nrm[3 * i + 1] = {t.n1.x, t.n1.y, t.n1.z, t.n.y};
        out.nrm[3 * i + 2] = {t.n2.x, t.n2.y, t.n2.z, t.n.z};
        out.nn[i] = {t.nn.x, t.nn.y, t.nn.z, 0.0f};
        if (!(t.alpha >= 1.0f)) out.opaque = 0;
        const int32_t texoff = t.texture_offset >= 0 ? t.texture_offset : -1;
        const int32_t tw = texoff >= 0 ? t.texture_width : 0, th = texoff >= 0 ? t.texture_height : 0;
        if (texoff >= 0) textured = true;
        const float vals[14] = {t.ar, t.ag, t.ab, t.alpha, t.er, t.eg, t.eb, t.ior, t.rough, t.metallic,
                                ibf(t.shader), ibf(texoff), ibf(tw), ibf(th)};
        MatKey key;
        std::memcpy(key.w, vals, sizeof(vals));
        auto it = mat_index.find(key);
        int32_t id;
        if (it == mat_index.end()) {
            id = int32_t(out.mats.size() / kMatWords);
            mat_index.emplace(key, id);
            out.mats.push_back({vals[0], vals[1], vals[2], vals[3]});
            out.mats.push_back({vals[4], vals[5], vals[6], vals[7]});
            out.mats.push_back({vals[8], vals[9], vals[10], vals[11]});
            out.mats.push_back({vals[12], vals[13], 0.0f, 0.0f});
        } else {
            id = it->second;
        }
        out.matid[i] = id;
    }
    if (textured) {
        out.uv.resize(s.ntri * 2);
        for (size_t i = 0; i < s.ntri; ++i) {
            const rm_triangle& t = s.tris[i];
            out.uv[2 * i] = {t.tu0, t.tv0, t.tu1, t.tv1};
            out.uv[2 * i + 1] = {t.tu2, t.tv2, 0.0f, 0.0f};
        }
    }
    // Lights in CDF order.
    out.lights.resize(s.nlight * kLightWords);
    out.cdf.assign(s.light_cdf, s.light_cdf + s.nlight);
    for (size_t k = 0; k < s.nlight; ++k) {
        const rm_triangle& t = s.tris[s.light_tri[k]];
        out.lights[4 * k] = {t.v0.x, t.v0.y, t.v0.z, t.v1.x};
        out.lights[4 * k + 1] = {t.v1.y, t.v1.z, t.v2.x, t.v2.y};
        out.lights[4 * k + 2] = {t.v2.z, t.nn.x, t.nn.y, t.nn.z};
        out.lights[4 * k + 3] = {t.er, t.eg, t.eb, t.alpha};
    }
    out.light_total = s.light_total;
    if (s.ntexel) out.texels.assign(s.texels, s.texels + 3 * s.ntexel);
    return true;
}

}  // namespace rmh
