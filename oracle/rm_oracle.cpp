// TEST INFRASTRUCTURE ONLY — the CPU oracle that the HIP backend is checked
// against. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg load it; the shipped library never includes, links or calls it.
//
// A plain C++ restatement of the reference's GPU path, one function per
// reference function, evaluated over the same Metal-ABI arrays the reference
// uploads:
//   Rng                         src/engine/pathtrace.metal:93-101
//   intersect / cast_ray        pathtrace.metal:119-204
//   samplers, BSDF              pathtrace.metal:65-91, 205-295
//   direct_lighting             pathtrace.metal:299-469
//   sample_dielectric           pathtrace.metal:470-534
//   cast_alpha_ray, pathtrace   pathtrace.metal:536-728
//   sample_sensor               pathtrace.metal:729-737
//   render_paths(_simd)         pathtrace.metal:739-815
//   restir_generate/spatial/shade pathtrace.metal:817-1258
//   denoise_temporal/atrous/tonemap pathtrace.metal:1263-1363
//   frame orchestration         src/engine/metal.mm:95-414
// Two deliberate definitions where the reference leaves results to traversal
// order or to its math library (both shared with the device code):
//   * closest hit = smallest distance, then smallest triangle index; box
//     culling keeps boxes whose entry lies within 2^-12 (relative) beyond the
//     current hit, so the answer no longer depends on the visit order;
//   * sin/cos of sampling angles use rm_sincosf (rm_sincos.h).
// Everything is templated on the real type R: R=float is the FP32 parity
// oracle, R=double the FP64 reference renderer.
#include "oracle_abi.h"
#include "../raymotion_amd/csrc/common/rm_atan.h"
#include "../raymotion_amd/csrc/common/rm_exp.h"
#include "../raymotion_amd/csrc/common/rm_fdiv.h"
#include "../raymotion_amd/csrc/common/rm_sincos.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <type_traits>
#include <thread>
#include <vector>

namespace orc {

// ---------------------------------------------------------------- helpers
struct Rng {  // PCG RXS-M-XS 32 seeded by splitmix64 (pathtrace.metal:93-101)
    uint32_t state = 0;
    void seed(uint64_t s) {
        uint64_t z = s + 0x9E3779B97F4A7C15ULL;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        state = uint32_t(z ^ (z >> 31));
    }
    float uniform01() {
        state = state * 747796405u + 2891336453u;
        const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
        return float(((word >> 22u) ^ word) >> 8u) * 0x1p-24f;
    }
};

template <class R> struct V3 {
    R x = 0, y = 0, z = 0;
    V3() = default;
    V3(R a, R b, R c) : x(a), y(b), z(c) {}
};
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator*(V3<R> a, V3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class R> inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline R dot3(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline V3<R> cross3(V3<R> a, V3<R> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> inline V3<R> vmin(V3<R> a, V3<R> b) { return {std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)}; }
template <class R> inline V3<R> vmax(V3<R> a, V3<R> b) { return {std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)}; }
template <class R> inline V3<R> cvt(const orc_vec3& v) { return {R(v.x), R(v.y), R(v.z)}; }
template <class R> inline R lum3(R r, R g, R b) { return R(0.2126) * r + R(0.7152) * g + R(0.0722) * b; }
template <class R> inline R fract(R x) { return std::fmin(x - std::floor(x), R(0.99999994)); }
template <class R> inline R gmod(R a, R b) { return a - std::floor(a / b) * b; }

template <class R> inline void sincos_(R x, R* s, R* c) {
    if constexpr (sizeof(R) == 4) {
        rm_sincosf(x, s, c);
    } else {
        *s = std::sin(x);
        *c = std::cos(x);
    }
}

template <class R> inline V3<R> normalize_(R x, R y, R z) {  // Metal normalize()
    const R inv = R(1) / std::sqrt(x * x + y * y + z * z);
    return {x * inv, y * inv, z * inv};
}

inline float tcull(float t) { return t + t * (1.0f / 4096.0f); }
inline double tcull(double t) { return t + t * (1.0 / 4096.0); }

template <class R> struct HitInfo {  // pathtrace.metal:11-23
    bool hit = false;
    R t = R(1e30);
    int tri = -1;
    R u = 0, v = 0;
    V3<R> p, n_interp;
    R ar = 0, ag = 0, ab = 0, er = 0, eg = 0, eb = 0, ior = 0, rough = 0;
    int shader = 0;
    V3<R> n_geom;
};

struct Stats { uint64_t nodes = 0, tris = 0, rays = 0; };

// A 4-wide traversal tree in the backend's layout (nodes: 8 float4 words
// each, the lower x, upper x, lower y, upper y, lower z and upper z bounds of
// the four children, then their references and kinds as int bits: -1 empty,
// 0 inner node, k > 0 leaf of k triangles from leaf-order index ref; geo: the
// triangles in leaf order, 3 float4 words each, v0.xyz v1.x | v1.yz v2.xy |
// v2.z id).
struct WideTree {
    const float* nodes = nullptr;
    size_t nnode = 0;
    const float* geo = nullptr;
};

// Set by orc_set_wide_tree: FP32 frames then traverse this tree with
// wide_walk in place of the binary walk of the scene (the same hits, and on
// the backend's own tree the same slab arithmetic, so also the same results
// where a ray grazes a box: tight boxes make such rays tree-dependent).
WideTree g_wide;

// The device's traversal (Tracer::intersect of the backend) over a wide
// tree: slab tests with fused multiply-adds, children sorted by entry
// distance, hit leaves intersected nearest-first within the current hit plus
// 2^-12, inner children descended nearest-first with the rest stacked; any:
// stop at the first hit. Returns the hit triangle (-1: none) and sets t (the
// hit distance, tmax without a hit), u, v; counts node visits and triangle
// tests into st.
inline int wide_walk(const WideTree& w, V3<float> ro, V3<float> rd, float tmax, bool any, Stats& st, float& t,
                     float& hu, float& hv) {
    auto ibits = [](float f) { int32_t i; std::memcpy(&i, &f, 4); return i; };
    float ht = tmax;
    int hid = -1;
    hu = hv = 0.0f;
    if (w.nnode > 0) {
        const float ix = 1.0f / (rd.x == 0.0f ? 1e-30f : rd.x), iy = 1.0f / (rd.y == 0.0f ? 1e-30f : rd.y),
                    iz = 1.0f / (rd.z == 0.0f ? 1e-30f : rd.z);
        const float ox = -ro.x * ix, oy = -ro.y * iy, oz = -ro.z * iz;
        std::vector<int> stack;
        int node = 0;
        bool done = false;
        while (!done) {
            const float* nw = w.nodes + size_t(node) * 32;
            ++st.nodes;
            const float tc = ht + ht * (1.0f / 4096.0f);
            float tn[4];
            int ref[4], kind[4];
            for (int c = 0; c < 4; ++c) {
                const float ax = std::fma(nw[0 + c], ix, ox), bx = std::fma(nw[4 + c], ix, ox);
                const float ay = std::fma(nw[8 + c], iy, oy), by = std::fma(nw[12 + c], iy, oy);
                const float az = std::fma(nw[16 + c], iz, oz), bz = std::fma(nw[20 + c], iz, oz);
                const float nr = std::fmax(std::fmax(std::fmin(ax, bx), std::fmin(ay, by)),
                                           std::fmax(std::fmin(az, bz), 0.0f));
                const float fr = std::fmin(std::fmin(std::fmax(ax, bx), std::fmax(ay, by)),
                                           std::fmin(std::fmax(az, bz), tc));
                const int k = ibits(nw[28 + c]);
                const bool ok = k >= 0 && nr <= fr;
                tn[c] = ok ? nr : INFINITY;
                ref[c] = ibits(nw[24 + c]);
                kind[c] = ok ? k : -1;
            }
            auto cx = [&](int a, int b) {
                if (tn[b] < tn[a]) { std::swap(tn[a], tn[b]); std::swap(ref[a], ref[b]); std::swap(kind[a], kind[b]); }
            };
            cx(0, 1); cx(2, 3); cx(0, 2); cx(1, 3); cx(1, 2);
            int next = -1;
            for (int c = 3; c >= 0; --c)
                if (kind[c] == 0) {
                    if (next >= 0) stack.push_back(next);
                    next = ref[c];
                }
            if (next < 0) {
                if (stack.empty()) done = true;
                else { next = stack.back(); stack.pop_back(); }
            }
            bool found = false;
            for (int c = 0; c < 4 && !found; ++c) {
                if (!(kind[c] > 0 && tn[c] <= ht + ht * (1.0f / 4096.0f))) continue;
                for (int k = 0; k < kind[c] && !found; ++k) {
                    const float* g = w.geo + size_t(ref[c] + k) * 12;
                    ++st.tris;
                    const V3<float> v0{g[0], g[1], g[2]}, v1{g[3], g[4], g[5]}, v2{g[6], g[7], g[8]};
                    const int id = ibits(g[9]);
                    const V3<float> e1 = v1 - v0, e2 = v2 - v0;
                    const V3<float> p = cross3(rd, e2);
                    const float det = dot3(e1, p);
                    if (std::fabs(det) < 1e-12f) continue;
                    const V3<float> tv = ro - v0;
                    const float u = dot3(tv, p) / det;
                    if (u < 0.0f || u > 1.0f) continue;
                    const V3<float> q = cross3(tv, e1);
                    const float v = dot3(rd, q) / det;
                    if (v < 0.0f || u + v > 1.0f) continue;
                    const float dd = dot3(e2, q) / det;
                    if (dd > 1e-8f && (dd < ht || (dd == ht && hid >= 0 && id < hid))) {
                        ht = dd;
                        hid = id;
                        hu = u;
                        hv = v;
                        if (any) found = true;
                    }
                }
            }
            if (found) break;
            node = next;
        }
    }
    t = ht;
    return hid;
}

// ---------------------------------------------------------------- tracer
template <class R> struct Tracer {
    const orc_scene* sc;
    orc_params P;
    Rng rng_;
    Stats* stats = nullptr;

    struct Bsdf { R r = 0, g = 0, b = 0, pdf = 0; };
    struct DielectricSample { bool valid = true; R n_x = 0, n_y = 0, n_z = 0, eta_scale = 1, dlw_da2 = 0; };
    // Derivatives of the opaque BSDF value: each channel by its own albedo
    // channel (da) and by alpha2 (dl). Not part of the reference.
    struct BsdfGrad { R da[3] = {0, 0, 0}, dl[3] = {0, 0, 0}; };
    // Legacy NEE sample that contributed: light triangle, its emission, the
    // BSDF value toward it, the factor k (contribution = le * e * k) and the
    // direction. Filled for the adjoint only.
    struct NeeOut { int tri = -1; R le[3] = {0, 0, 0}, e[3] = {0, 0, 0}; R k = 0; V3<R> dir; };

    R urand() { return R(rng_.uniform01()); }

    // pathtrace.metal:119-163 (software branch) with the canonical tie-break.
    HitInfo<R> intersect(V3<R> ro, V3<R> rd, R dist, bool shadow) {
        HitInfo<R> hit;
        hit.t = dist;
        if (stats) ++stats->rays;
        if constexpr (std::is_same<R, float>::value) {
            if (g_wide.nodes) {
                Stats local;
                float t, u, v;
                hit.tri = wide_walk(g_wide, ro, rd, dist, shadow, stats ? *stats : local, t, u, v);
                hit.t = t;
                hit.hit = hit.tri >= 0;
                hit.u = u;
                hit.v = v;
                hit.p = ro + rd * hit.t;
                return hit;
            }
        }
        if (P.nodes == 0) return hit;
        const V3<R> safe{rd.x == 0 ? R(1e-30) : rd.x, rd.y == 0 ? R(1e-30) : rd.y, rd.z == 0 ? R(1e-30) : rd.z};
        const V3<R> inv{R(1) / safe.x, R(1) / safe.y, R(1) / safe.z};
        int ni = 0;
        while (ni < P.nodes) {
            const orc_node& n = sc->nodes[ni];
            if (stats) ++stats->nodes;
            const V3<R> a = (cvt<R>(n.mn) - ro) * inv, b = (cvt<R>(n.mx) - ro) * inv;
            const V3<R> lo = vmin(a, b), hi = vmax(a, b);
            const R near_t = std::fmax(R(0), std::fmax(lo.x, std::fmax(lo.y, lo.z)));
            const R far_t = std::fmin(tcull(hit.t), std::fmin(hi.x, std::fmin(hi.y, hi.z)));
            if (near_t > far_t) { ni = n.right; continue; }
            ++ni;
            if (n.count == 0) continue;
            for (int k = 0; k < n.count; ++k) {
                const int j = sc->indices[n.offset + k];
                const orc_triangle& t = sc->tris[j];
                if (stats) ++stats->tris;
                const V3<R> v0 = cvt<R>(t.v0);
                const V3<R> e1 = cvt<R>(t.v1) - v0, e2 = cvt<R>(t.v2) - v0;
                const V3<R> p = cross3(rd, e2);
                const R det = dot3(e1, p);
                if (std::fabs(det) < R(1e-12)) continue;
                const V3<R> tv = ro - v0;
                const R u = dot3(tv, p) / det;
                if (u < 0 || u > 1) continue;
                const V3<R> q = cross3(tv, e1);
                const R v = dot3(rd, q) / det;
                if (v < 0 || u + v > 1) continue;
                const R d = dot3(e2, q) / det;
                if (d > R(1e-8) && (d < hit.t || (d == hit.t && hit.tri >= 0 && j < hit.tri))) {
                    hit.t = d; hit.tri = j; hit.u = u; hit.v = v; hit.hit = true;
                    if (shadow) return hit;
                }
            }
        }
        hit.p = ro + rd * hit.t;
        return hit;
    }
    bool cast_shadow_ray(V3<R> ro, V3<R> rd, R dist) { return intersect(ro, rd, dist, true).hit; }

    HitInfo<R> cast_ray(V3<R> ro, V3<R> rd, R dist) {  // pathtrace.metal:165-204
        HitInfo<R> hit = intersect(ro, rd, dist, false);
        if (hit.tri >= 0) {
            const orc_triangle& t = sc->tris[hit.tri];
            const R w = 1 - (hit.u + hit.v);
            hit.n_interp = normalize_(w * R(t.n0.x) + (hit.u * R(t.n1.x) + hit.v * R(t.n2.x)),
                                      w * R(t.n0.y) + (hit.u * R(t.n1.y) + hit.v * R(t.n2.y)),
                                      w * R(t.n0.z) + (hit.u * R(t.n1.z) + hit.v * R(t.n2.z)));
            hit.ar = t.ar; hit.ag = t.ag; hit.ab = t.ab;
            if (t.texture_offset >= 0) {
                const R u = fract(w * R(t.tu0) + hit.u * R(t.tu1) + hit.v * R(t.tu2));
                const R v = fract(w * R(t.tv0) + hit.u * R(t.tv1) + hit.v * R(t.tv2));
                const int x = std::min(int(u * R(t.texture_width)), t.texture_width - 1);
                const int y = t.texture_height - 1 - std::min(int(v * R(t.texture_height)), t.texture_height - 1);
                const float* c = sc->texels + 3 * size_t(t.texture_offset + y * t.texture_width + x);
                hit.ar *= R(c[0]); hit.ag *= R(c[1]); hit.ab *= R(c[2]);
            }
            hit.er = t.er; hit.eg = t.eg; hit.eb = t.eb;
            hit.ior = t.ior; hit.rough = t.rough; hit.shader = t.shader;
            hit.n_geom = cvt<R>(t.n);
            if (hit.shader == 1) {
                const R kpi = R(3.141592653589793);
                const R az = atan_r(hit.p.x / hit.p.z) * R(180) / kpi + ((hit.p.z > 0) ? R(0) : R(180));
                const R h = gmod(az, R(360)) / R(360);
                // hsv2rgb(h, 0.8, 1) (pathtrace.metal:27-64); h lies in [0,1) here.
                const R s = R(0.8), val = R(1);
                const R c = s * val;
                const R x = c * (1 - std::fabs(gmod(h / R(60), R(2)) - 1));
                const R m = val - c;
                R cr = 0, cg = 0, cb = 0;
                if (h < 60) { cr = c; cg = x; }
                else if (h < 120) { cr = x; cg = c; }
                else if (h < 180) { cg = c; cb = x; }
                else if (h < 240) { cg = x; cb = c; }
                else if (h < 300) { cr = x; cb = c; }
                else if (h < 360) { cr = c; cb = x; }
                hit.ar = (cr + m) * 255; hit.ag = (cg + m) * 255; hit.ab = (cb + m) * 255;
            }
        }
        return hit;
    }

    static void make_tangent(R nx, R ny, R nz, R& tx, R& ty, R& tz, R& bx, R& by, R& bz) {
        if (std::fabs(nx) > R(0.1)) { tx = nz; ty = 0; tz = -nx; }
        else { tx = 0; ty = -nz; tz = ny; }
        const R l = R(1) / std::sqrt(tx * tx + ty * ty + tz * tz);
        tx *= l; ty *= l; tz *= l;
        bx = ny * tz - nz * ty;
        by = nz * tx - nx * tz;
        bz = nx * ty - ny * tx;
    }

    void sample_cosine_hemisphere(R nx, R ny, R nz, R& ox, R& oy, R& oz) {
        const R ru = urand();
        const R phi = urand() * R(6.2831855f);
        R sp, c;
        sincos_(phi, &sp, &c);
        const R rr = std::sqrt(ru);
        const R lx = rr * c, ly = rr * sp, lz = std::sqrt(std::fmax(R(0), 1 - ru));
        R tx, ty, tz, bx, by, bz;
        make_tangent(nx, ny, nz, tx, ty, tz, bx, by, bz);
        ox = tx * lx + (bx * ly + nx * lz);
        oy = ty * lx + (by * ly + ny * lz);
        oz = tz * lx + (bz * ly + nz * lz);
    }

    void sample_ggx_half(R alpha2, R nx, R ny, R nz, R& hx, R& hy, R& hz) {
        (void)urand();  // the first draw is unused (pathtrace.metal:224)
        const R rv = urand();
        const R phi = urand() * R(6.2831855f);
        R sp, c;
        sincos_(phi, &sp, &c);
        const R cs = std::sqrt((1 - rv) / (1 + (alpha2 - 1) * rv));
        const R sn = std::sqrt(std::fmax(R(0), 1 - cs * cs));
        const R lx = sn * c, ly = sn * sp, lz = cs;
        R tx, ty, tz, bx, by, bz;
        make_tangent(nx, ny, nz, tx, ty, tz, bx, by, bz);
        hx = tx * lx + (bx * ly + nx * lz);
        hy = ty * lx + (by * ly + ny * lz);
        hz = tz * lx + (bz * ly + nz * lz);
        const R d = R(1) / std::sqrt(hx * hx + hy * hy + hz * hz);
        hx *= d; hy *= d; hz *= d;
    }

    // x / pi: in FP32 the two-term-reciprocal division both sides share
    // (rm_fdiv.h; correctly rounded for 2^-100 <= |x| < 2^100), in FP64 the
    // IEEE division.
    static R div_pi(R x) {
        if constexpr (std::is_same<R, float>::value) return rm_div_pi(x);
        else return x / R(3.141592653589793);
    }
    // atan of the hue shader: in FP32 the deterministic rm_atanf (rm_atan.h,
    // within one ulp of libm, the same bits as the kernels), in FP64 libm's.
    static R atan_r(R x) {
        if constexpr (std::is_same<R, float>::value) return rm_atanf(x);
        else return std::atan(x);
    }

    static Bsdf evaluate_opaque_bsdf(R wi_x, R wi_y, R wi_z, R wo_x, R wo_y, R wo_z, R n_x, R n_y, R n_z,
                              R alpha2, R metallic, R spec_prob, R ar, R ag, R ab) {
        const R kpi = R(3.141592653589793);
        Bsdf out;
        const R cos_i = wi_x * n_x + (wi_y * n_y + wi_z * n_z);
        const R cos_o = wo_x * n_x + (wo_y * n_y + wo_z * n_z);
        if (!(cos_i > R(1e-8) && cos_o > R(1e-8))) return out;
        R hx = wi_x + wo_x, hy = wi_y + wo_y, hz = wi_z + wo_z;
        const R hl = std::sqrt(hx * hx + hy * hy + hz * hz);
        if (!(hl > R(1e-8))) return out;
        const R inv = R(1) / hl;
        hx *= inv; hy *= inv; hz *= inv;
        R cos_h = hx * n_x + (hy * n_y + hz * n_z);
        R vo_h = wo_x * hx + (wo_y * hy + wo_z * hz);
        if (cos_h < 0) { cos_h = -cos_h; vo_h = -vo_h; }
        if (!(vo_h > R(1e-8))) return out;
        const R dd = cos_h * cos_h * (alpha2 - 1) + 1;
        const R D = alpha2 / (kpi * dd * dd);
        const R g1 = 2 * cos_o / (cos_o + std::sqrt(cos_o * cos_o + alpha2 * (1 - cos_o * cos_o)));
        const R g2 = 2 * cos_i / (cos_i + std::sqrt(cos_i * cos_i + alpha2 * (1 - cos_i * cos_i)));
        const R G = g1 * g2;
        R om = 1 - vo_h;
        om *= om;
        om *= om;
        om *= 1 - vo_h;
        const R di = 1 - metallic;
        const R f0r = R(0.04) * di + ar * metallic, f0g = R(0.04) * di + ag * metallic, f0b = R(0.04) * di + ab * metallic;
        const R fr = f0r + (1 - f0r) * om, fg = f0g + (1 - f0g) * om, fb = f0b + (1 - f0b) * om;
        const R kdr = ar * di * (1 - fr), kdg = ag * di * (1 - fg), kdb = ab * di * (1 - fb);
        const R base = D * G / (4 * cos_i * cos_o);
        out.r = div_pi(kdr) + base * fr;
        out.g = div_pi(kdg) + base * fg;
        out.b = div_pi(kdb) + base * fb;
        const R pdf_d = div_pi(cos_i);
        const R pdf_s = D * cos_h / (4 * vo_h);
        out.pdf = (1 - spec_prob) * pdf_d + spec_prob * pdf_s;
        return out;
    }

    // d f / d albedo and d f / d alpha2 of evaluate_opaque_bsdf from the
    // textbook derivatives: Schlick F' = m (1 - om), GGX
    // D' = (dd - 2 alpha2 cos_h^2) / (pi dd^3), Smith G1' = -c (1 - c^2) / (s (c + s)^2),
    // s = sqrt(c^2 + alpha2 (1 - c^2)). The pdf is not differentiated.
    static BsdfGrad evaluate_opaque_bsdf_grad(R wi_x, R wi_y, R wi_z, R wo_x, R wo_y, R wo_z, R n_x, R n_y, R n_z,
                                              R alpha2, R metallic, R ar, R ag, R ab) {
        const R kpi = R(3.141592653589793);
        BsdfGrad out;
        const R cos_i = wi_x * n_x + (wi_y * n_y + wi_z * n_z);
        const R cos_o = wo_x * n_x + (wo_y * n_y + wo_z * n_z);
        if (!(cos_i > R(1e-8) && cos_o > R(1e-8))) return out;
        R hx = wi_x + wo_x, hy = wi_y + wo_y, hz = wi_z + wo_z;
        const R hl = std::sqrt(hx * hx + hy * hy + hz * hz);
        if (!(hl > R(1e-8))) return out;
        const R inv = R(1) / hl;
        hx *= inv; hy *= inv; hz *= inv;
        R cos_h = hx * n_x + (hy * n_y + hz * n_z);
        R vo_h = wo_x * hx + (wo_y * hy + wo_z * hz);
        if (cos_h < 0) { cos_h = -cos_h; vo_h = -vo_h; }
        if (!(vo_h > R(1e-8))) return out;
        const R dd = cos_h * cos_h * (alpha2 - 1) + 1;
        const R D = alpha2 / (kpi * dd * dd);
        const R so = std::sqrt(cos_o * cos_o + alpha2 * (1 - cos_o * cos_o));
        const R si = std::sqrt(cos_i * cos_i + alpha2 * (1 - cos_i * cos_i));
        const R g1 = 2 * cos_o / (cos_o + so);
        const R g2 = 2 * cos_i / (cos_i + si);
        const R G = g1 * g2;
        R om = 1 - vo_h;
        om *= om;
        om *= om;
        om *= 1 - vo_h;
        const R di = 1 - metallic;
        const R a[3] = {ar, ag, ab};
        const R base = D * G / (4 * cos_i * cos_o);
        const R dF = metallic * (1 - om);
        const R dD = (dd - 2 * alpha2 * cos_h * cos_h) / (kpi * dd * dd * dd);
        const R dg1 = -cos_o * (1 - cos_o * cos_o) / (so * (cos_o + so) * (cos_o + so));
        const R dg2 = -cos_i * (1 - cos_i * cos_i) / (si * (cos_i + si) * (cos_i + si));
        const R dbase = (dD * G + D * (dg1 * g2 + g1 * dg2)) / (4 * cos_i * cos_o);
        for (int c = 0; c < 3; ++c) {
            const R f0 = R(0.04) * di + a[c] * metallic;
            const R F = f0 + (1 - f0) * om;
            out.da[c] = div_pi(di * (1 - F) - a[c] * di * dF) + base * dF;
            out.dl[c] = dbase * F;
        }
        return out;
    }

    int pick_light(R pick) {  // binary search over the CDF (pathtrace.metal:323-330)
        int lo = 0, hi = P.lights - 1;
        while (lo < hi) {
            const int mid = (lo + hi) / 2;
            if (R(sc->cdf[mid]) >= pick) hi = mid;
            else lo = mid + 1;
        }
        return sc->lights[lo];
    }

    bool occluded(V3<R> o, V3<R> d, R dist) {
        return P.opaque ? cast_shadow_ray(o, d, dist) : cast_alpha_ray(o, d, dist).tri >= 0;
    }

    void direct_lighting(R p_x, R p_y, R p_z, R n_x, R n_y, R n_z, R wo_x, R wo_y, R wo_z, R rough,
                         R metallic, R spec_prob, R ar, R ag, R ab, R& or_, R& og, R& ob, R& out_sumW,
                         uint32_t& out_M, NeeOut* no = nullptr) {
        or_ = og = ob = 0;
        out_sumW = 0;
        out_M = 0;
        if (P.lights == 0 || !(R(P.light_total) > R(1e-12))) return;
        int Mc = P.restir_candidates;
        if (Mc < 1) Mc = 1;
        if (Mc > 32) Mc = 32;
        const R sp = spec_prob;
        R r2 = rough;
        if (r2 < 0) r2 = 0;
        if (r2 > 1) r2 = 1;
        if (r2 < R(0.02)) r2 = R(0.02);
        const R alpha2 = (r2 * r2) * (r2 * r2);
        const bool nomis = P.nomis != 0;
        const R total = R(P.light_total);
        if (Mc <= 1) {  // legacy single-sample NEE (pathtrace.metal:319-384)
            const R pick = urand() * total;
            const int light_tri = pick_light(pick);
            const orc_triangle& t = sc->tris[light_tri];
            R lu = urand(), lv = urand();
            if (lu + lv > 1) { lu = 1 - lu; lv = 1 - lv; }
            const R lw = 1 - lu - lv;
            R ldx = lw * R(t.v0.x) + (lu * R(t.v1.x) + lv * R(t.v2.x)) - p_x;
            R ldy = lw * R(t.v0.y) + (lu * R(t.v1.y) + lv * R(t.v2.y)) - p_y;
            R ldz = lw * R(t.v0.z) + (lu * R(t.v1.z) + lv * R(t.v2.z)) - p_z;
            const R dist = std::sqrt(ldx * ldx + ldy * ldy + ldz * ldz);
            if (!(dist > R(0.003))) { out_M = 1; return; }
            const R inv = R(1) / dist;
            ldx *= inv; ldy *= inv; ldz *= inv;
            const R lcos = -(R(t.nn.x) * ldx + (R(t.nn.y) * ldy + R(t.nn.z) * ldz));
            const R scos = n_x * ldx + (n_y * ldy + n_z * ldz);
            if (!(lcos > R(1e-8) && scos > R(1e-8))) { out_M = 1; return; }
            R lt_lum = lum3<R>(t.er, t.eg, t.eb);
            if (lt_lum < R(1e-6)) lt_lum = R(1e-6);
            const R pdf = dist * dist * lt_lum / (lcos * total);
            const V3<R> so{p_x + n_x * R(0.002), p_y + n_y * R(0.002), p_z + n_z * R(0.002)};
            if (occluded(so, {ldx, ldy, ldz}, dist - R(0.003))) {
                const Bsdf eo = evaluate_opaque_bsdf(ldx, ldy, ldz, wo_x, wo_y, wo_z, n_x, n_y, n_z, alpha2,
                                                     metallic, sp, ar, ag, ab);
                const R ph0 = lum3<R>(R(t.er) * eo.r * scos * R(t.alpha), R(t.eg) * eo.g * scos * R(t.alpha),
                                      R(t.eb) * eo.b * scos * R(t.alpha));
                out_sumW = (ph0 > 0 && pdf > R(1e-12)) ? ph0 / pdf : R(0);
                out_M = 1;
                return;
            }
            const Bsdf e = evaluate_opaque_bsdf(ldx, ldy, ldz, wo_x, wo_y, wo_z, n_x, n_y, n_z, alpha2,
                                                metallic, sp, ar, ag, ab);
            if (!(pdf > R(1e-12))) { out_M = 1; return; }
            const R Fr = R(t.er) * e.r * scos * R(t.alpha), Fg = R(t.eg) * e.g * scos * R(t.alpha),
                    Fb = R(t.eb) * e.b * scos * R(t.alpha);
            const R ph = lum3(Fr, Fg, Fb);
            out_sumW = (ph > 0) ? ph / pdf : R(0);
            out_M = 1;
            R wnee = 1;
            if (!nomis && e.pdf > R(1e-20)) wnee = pdf / (pdf + e.pdf);
            const R k = scos * wnee * R(t.alpha) / pdf;
            or_ = R(t.er) * e.r * k;
            og = R(t.eg) * e.g * k;
            ob = R(t.eb) * e.b * k;
            if (no) {
                no->tri = light_tri;
                no->le[0] = R(t.er); no->le[1] = R(t.eg); no->le[2] = R(t.eb);
                no->e[0] = e.r; no->e[1] = e.g; no->e[2] = e.b;
                no->k = k;
                no->dir = {ldx, ldy, ldz};
            }
            return;
        }
        // RIS over Mc candidates, one shadow ray (pathtrace.metal:385-469).
        R sumW = 0;
        R sel_lx = 0, sel_ly = 0, sel_lz = 0, sel_Fr = 0, sel_Fg = 0, sel_Fb = 0, sel_phat = 0, sel_epdf = 0;
        bool have = false;
        for (int j = 0; j < Mc; ++j) {
            const R pick = urand() * total;
            const orc_triangle& t = sc->tris[pick_light(pick)];
            R lu = urand(), lv = urand();
            if (lu + lv > 1) { lu = 1 - lu; lv = 1 - lv; }
            const R lw = 1 - lu - lv;
            const R qx = lw * R(t.v0.x) + (lu * R(t.v1.x) + lv * R(t.v2.x));
            const R qy = lw * R(t.v0.y) + (lu * R(t.v1.y) + lv * R(t.v2.y));
            const R qz = lw * R(t.v0.z) + (lu * R(t.v1.z) + lv * R(t.v2.z));
            const R ldx = qx - p_x, ldy = qy - p_y, ldz = qz - p_z;
            const R dist = std::sqrt(ldx * ldx + ldy * ldy + ldz * ldz);
            R w = 0, ph = 0, epdf = 0, Fr = 0, Fg = 0, Fb = 0;
            if (dist > R(0.003)) {
                const R inv = R(1) / dist;
                const R dxn = ldx * inv, dyn = ldy * inv, dzn = ldz * inv;
                const R lcos = -(R(t.nn.x) * dxn + (R(t.nn.y) * dyn + R(t.nn.z) * dzn));
                const R scos = n_x * dxn + (n_y * dyn + n_z * dzn);
                if (lcos > R(1e-8) && scos > R(1e-8)) {
                    R lt_lum = lum3<R>(t.er, t.eg, t.eb);
                    if (lt_lum < R(1e-6)) lt_lum = R(1e-6);
                    const R q = dist * dist * lt_lum / (lcos * total);
                    if (q > R(1e-12)) {
                        const Bsdf e = evaluate_opaque_bsdf(dxn, dyn, dzn, wo_x, wo_y, wo_z, n_x, n_y, n_z, alpha2,
                                                            metallic, sp, ar, ag, ab);
                        epdf = e.pdf;
                        Fr = R(t.er) * e.r * scos * R(t.alpha);
                        Fg = R(t.eg) * e.g * scos * R(t.alpha);
                        Fb = R(t.eb) * e.b * scos * R(t.alpha);
                        ph = lum3(Fr, Fg, Fb);
                        if (ph > 0) w = ph / q;
                    }
                }
            }
            sumW += w;
            const R r = urand();
            if (w > 0 && r * sumW < w) {
                have = true;
                sel_lx = qx; sel_ly = qy; sel_lz = qz;
                sel_Fr = Fr; sel_Fg = Fg; sel_Fb = Fb;
                sel_phat = ph; sel_epdf = epdf;
            }
        }
        out_sumW = sumW;
        out_M = uint32_t(Mc);
        if (!have || !(sumW > 0) || !(sel_phat > 0)) return;
        const R pdf_ris = R(Mc) * sel_phat / sumW;
        R wnee = 1;
        if (!nomis && sel_epdf > R(1e-20)) wnee = pdf_ris / (pdf_ris + sel_epdf);
        R vdx = sel_lx - p_x, vdy = sel_ly - p_y, vdz = sel_lz - p_z;
        const R vdist = std::sqrt(vdx * vdx + vdy * vdy + vdz * vdz);
        if (!(vdist > R(0.003))) return;
        const R vinv = R(1) / vdist;
        vdx *= vinv; vdy *= vinv; vdz *= vinv;
        const V3<R> so{p_x + n_x * R(0.002), p_y + n_y * R(0.002), p_z + n_z * R(0.002)};
        if (occluded(so, {vdx, vdy, vdz}, vdist - R(0.003))) return;
        const R W = sumW / (R(Mc) * sel_phat);
        or_ = sel_Fr * W * wnee;
        og = sel_Fg * W * wnee;
        ob = sel_Fb * W * wnee;
    }

    static R dielectric_fresnel(R ci, R eta_in, R eta_out) {  // pathtrace.metal:80-91
        if (ci < 0) ci = 0;
        if (ci > 1) ci = 1;
        const R eta = eta_in / eta_out;
        const R s2 = eta * eta * (1 - ci * ci);
        if (s2 >= 1) return 1;
        const R ct = std::sqrt(1 - s2);
        const R rp = (eta_out * ci - eta_in * ct) / (eta_out * ci + eta_in * ct);
        const R rs = (eta_in * ci - eta_out * ct) / (eta_in * ci + eta_out * ct);
        return R(0.5) * (rp * rp + rs * rs);
    }

    DielectricSample sample_dielectric(R wo_x, R wo_y, R wo_z, R n_x, R n_y, R n_z, R curr_dx, R curr_dy,
                                       R curr_dz, R ior, bool front_face, R rough, R alpha2) {
        DielectricSample out;
        R eta_i = 1, eta_t = ior;
        if (!front_face) { eta_i = ior; eta_t = 1; }
        R hx, hy, hz;
        if (rough <= R(0.0201)) {
            hx = n_x; hy = n_y; hz = n_z;
        } else {
            sample_ggx_half(alpha2, n_x, n_y, n_z, hx, hy, hz);
            const R md = wo_x * hx + (wo_y * hy + wo_z * hz);
            if (!(md > R(1e-8))) out.valid = false;
        }
        R cos_m = wo_x * hx + (wo_y * hy + wo_z * hz);
        if (cos_m < 0) { hx = -hx; hy = -hy; hz = -hz; cos_m = -cos_m; }
        if (cos_m > 1) cos_m = 1;
        const R fres = dielectric_fresnel(cos_m, eta_i, eta_t);
        const R eta = eta_i / eta_t;
        const R s2 = eta * eta * (1 - cos_m * cos_m);
        if (s2 >= 1 || urand() < fres) {
            out.n_x = 2 * cos_m * hx - wo_x;
            out.n_y = 2 * cos_m * hy - wo_y;
            out.n_z = 2 * cos_m * hz - wo_z;
            out.eta_scale = 1;
        } else {
            const R cos_t = std::sqrt(1 - s2);
            out.n_x = eta * curr_dx + hx * (eta * cos_m - cos_t);
            out.n_y = eta * curr_dy + hy * (eta * cos_m - cos_t);
            out.n_z = eta * curr_dz + hz * (eta * cos_m - cos_t);
            out.eta_scale = eta * eta;
        }
        {
            const R d = R(1) / std::sqrt(out.n_x * out.n_x + out.n_y * out.n_y + out.n_z * out.n_z);
            out.n_x *= d; out.n_y *= d; out.n_z *= d;
        }
        if (rough > R(0.0201)) {
            const R cos_i = std::fabs(out.n_x * n_x + (out.n_y * n_y + out.n_z * n_z));
            const R cos_o = std::fabs(wo_x * n_x + (wo_y * n_y + wo_z * n_z));
            const R cos_h = std::fabs(hx * n_x + (hy * n_y + hz * n_z));
            const R vo_h = std::fabs(wo_x * hx + (wo_y * hy + wo_z * hz));
            if (cos_i > R(1e-8) && cos_o > R(1e-8) && cos_h > R(1e-8) && vo_h > R(1e-8)) {
                const R so = std::sqrt(cos_o * cos_o + alpha2 * (1 - cos_o * cos_o));
                const R si = std::sqrt(cos_i * cos_i + alpha2 * (1 - cos_i * cos_i));
                const R g1 = 2 * cos_o / (cos_o + so);
                const R g2 = 2 * cos_i / (cos_i + si);
                R w = g1 * g2 * vo_h / (cos_o * cos_h);
                if (!(w > 10))  // d log(G1 G2) / d alpha2, zero where the weight is clamped
                    out.dlw_da2 = -(1 - cos_o * cos_o) / (2 * so * (cos_o + so)) -
                                  (1 - cos_i * cos_i) / (2 * si * (cos_i + si));
                if (w > 10) w = 10;
                out.eta_scale *= w;
            } else {
                out.valid = false;
            }
        }
        return out;
    }

    HitInfo<R> cast_alpha_ray(V3<R> ro, V3<R> rd, R dist) {  // pathtrace.metal:536-550
        if (P.opaque) return cast_ray(ro, rd, dist);
        R traveled = 0;
        while (traveled < dist) {
            HitInfo<R> hit = cast_ray(ro + rd * traveled, rd, dist - traveled);
            if (hit.tri < 0) return hit;
            const R opacity = R(sc->tris[hit.tri].alpha);
            if (opacity >= 1 || (opacity > 0 && urand() < opacity)) {
                hit.t += traveled;
                return hit;
            }
            traveled += hit.t + R(0.0001);
        }
        return HitInfo<R>{};
    }

    // Shared "emission seen along a BSDF-sampled ray" MIS weight
    // (pathtrace.metal:595-636 and 1191-1209).
    R emission_mis(const HitInfo<R>& hit, R ox, R oy, R oz, R last_pdf, R prev_sumW, uint32_t prev_M, R pb_r,
                   R pb_g, R pb_b, R pb_cos) {
        const orc_triangle& lt = sc->tris[hit.tri];
        R lx = hit.p.x - ox, ly = hit.p.y - oy, lz = hit.p.z - oz;
        const R dist2 = lx * lx + ly * ly + lz * lz;
        const R dist = std::sqrt(dist2);
        R pl = R(1e30);
        if (dist > R(1e-20) && last_pdf > R(1e-20) && prev_M > 0 && prev_sumW > R(1e-20)) {
            const R ph = lum3(hit.er * pb_r * pb_cos * R(lt.alpha), hit.eg * pb_g * pb_cos * R(lt.alpha),
                              hit.eb * pb_b * pb_cos * R(lt.alpha));
            if (ph > R(1e-20)) {
                pl = R(prev_M) * ph / prev_sumW;
            } else {
                const R inv = R(1) / dist;
                const R lxn = lx * inv, lyn = ly * inv, lzn = lz * inv;
                const R cos_l = -(R(lt.nn.x) * lxn + (R(lt.nn.y) * lyn + R(lt.nn.z) * lzn));
                if (cos_l > R(1e-8)) {
                    R lum = lum3<R>(lt.er, lt.eg, lt.eb);
                    if (lum < R(1e-6)) lum = R(1e-6);
                    pl = dist2 * lum / (cos_l * R(P.light_total));
                }
            }
        } else if (dist > R(1e-20) && last_pdf > R(1e-20)) {
            const R inv = R(1) / dist;
            lx *= inv; ly *= inv; lz *= inv;
            const R cos_l = -(R(lt.nn.x) * lx + (R(lt.nn.y) * ly + R(lt.nn.z) * lz));
            if (cos_l > R(1e-8)) {
                R lum = lum3<R>(lt.er, lt.eg, lt.eb);
                if (lum < R(1e-6)) lum = R(1e-6);
                pl = dist2 * lum / (cos_l * R(P.light_total));
            }
        }
        return P.nomis ? R(0) : ((last_pdf + pl > R(1e-20)) ? last_pdf / (last_pdf + pl) : R(0));
    }

    struct NoAdjoint {
        template <class... T> void emission(const T&...) {}
        template <class... T> void nee(const T&...) {}
        template <class... T> void scatter(const T&...) {}
        template <class... T> void dielectric(const T&...) {}
    };
    V3<R> pathtrace(V3<R> ro, V3<R> rd, int max_bounces) {
        NoAdjoint none;
        return pathtrace_t(ro, rd, max_bounces, none);
    }
    // pathtrace.metal:552-728; the hooks of `adj` see each term of the estimate.
    template <class A> V3<R> pathtrace_t(V3<R> ro, V3<R> rd, int max_bounces, A& adj) {
        R ox = ro.x, oy = ro.y, oz = ro.z, dx = rd.x, dy = rd.y, dz = rd.z;
        R accr = 0, accg = 0, accb = 0, thr_r = 1, thr_g = 1, thr_b = 1;
        int bounce = 0, prev_delta = 1;
        R last_pdf = 0, prev_sumW = 0;
        uint32_t prev_M = 0;
        R prev_br = 0, prev_bg = 0, prev_bb = 0, prev_cos = 0;
        for (int b = 0; b < max_bounces; ++b) {
            const HitInfo<R> hit = cast_alpha_ray({ox, oy, oz}, {dx, dy, dz}, R(1e30));
            if (hit.tri < 0) break;
            R pnx = hit.n_interp.x, pny = hit.n_interp.y, pnz = hit.n_interp.z;
            const R geom = hit.n_geom.x * pnx + (hit.n_geom.y * pny + hit.n_geom.z * pnz);
            if (geom < 0) { pnx = -pnx; pny = -pny; pnz = -pnz; }
            const R ndot = dx * pnx + (dy * pny + dz * pnz);
            bool front = true;
            if (ndot > 0) { pnx = -pnx; pny = -pny; pnz = -pnz; front = false; }
            const R wox = -dx, woy = -dy, woz = -dz;
            R rough = hit.rough;
            if (rough < 0) rough = 0;
            if (rough > 1) rough = 1;
            if (rough < R(0.02)) rough = R(0.02);
            const R alpha = rough * rough;
            const R alpha2 = alpha * alpha;
            R metallic = R(sc->tris[hit.tri].metallic);
            if (metallic < 0) metallic = 0;
            if (metallic > 1) metallic = 1;
            if (bounce == 0 || prev_delta == 1) {
                accr += hit.er * thr_r; accg += hit.eg * thr_g; accb += hit.eb * thr_b;
                adj.emission(hit, thr_r, thr_g, thr_b, R(1));
            } else {
                const R w = emission_mis(hit, ox, oy, oz, last_pdf, prev_sumW, prev_M, prev_br, prev_bg, prev_bb,
                                         prev_cos);
                accr += hit.er * thr_r * w; accg += hit.eg * thr_g * w; accb += hit.eb * thr_b * w;
                adj.emission(hit, thr_r, thr_g, thr_b, w);
            }
            if (hit.ior > R(1.0001)) {
                const DielectricSample ds =
                    sample_dielectric(wox, woy, woz, pnx, pny, pnz, dx, dy, dz, hit.ior, front, rough, alpha2);
                if (!ds.valid) break;
                adj.dielectric(hit, ds, accr, accg, accb);
                thr_r *= ds.eta_scale; thr_g *= ds.eta_scale; thr_b *= ds.eta_scale;
                R s = 1;
                if (ds.n_x * pnx + (ds.n_y * pny + ds.n_z * pnz) < 0) s = -1;
                ox = hit.p.x + pnx * R(0.001) * s;
                oy = hit.p.y + pny * R(0.001) * s;
                oz = hit.p.z + pnz * R(0.001) * s;
                dx = ds.n_x; dy = ds.n_y; dz = ds.n_z;
                prev_delta = 1;
                prev_sumW = 0; prev_M = 0;
            } else {
                R spec_prob = R(0.05) + R(0.95) * metallic;
                if (spec_prob > R(0.98)) spec_prob = R(0.98);
                if (metallic >= R(0.999)) spec_prob = 1;
                R cur_sumW = 0;
                uint32_t cur_M = 0;
                {
                    R dlr, dlg, dlb;
                    NeeOut no;
                    direct_lighting(hit.p.x, hit.p.y, hit.p.z, pnx, pny, pnz, wox, woy, woz, rough, metallic,
                                    spec_prob, hit.ar, hit.ag, hit.ab, dlr, dlg, dlb, cur_sumW, cur_M, &no);
                    accr += dlr * thr_r; accg += dlg * thr_g; accb += dlb * thr_b;
                    if (no.tri >= 0) adj.nee(hit, no, thr_r, thr_g, thr_b, wox, woy, woz, pnx, pny, pnz, alpha2, metallic);
                }
                R nx2, ny2, nz2;
                if (urand() < spec_prob) {
                    R hx, hy, hz;
                    sample_ggx_half(alpha2, pnx, pny, pnz, hx, hy, hz);
                    const R md = wox * hx + (woy * hy + woz * hz);
                    if (!(md > R(1e-8))) break;
                    nx2 = 2 * md * hx - wox; ny2 = 2 * md * hy - woy; nz2 = 2 * md * hz - woz;
                    prev_delta = (rough <= R(0.0201)) ? 1 : 0;
                } else {
                    sample_cosine_hemisphere(pnx, pny, pnz, nx2, ny2, nz2);
                    prev_delta = 0;
                }
                const R ncos = nx2 * pnx + (ny2 * pny + nz2 * pnz);
                if (!(ncos > R(1e-8))) break;
                {
                    const R d = R(1) / std::sqrt(nx2 * nx2 + ny2 * ny2 + nz2 * nz2);
                    nx2 *= d; ny2 *= d; nz2 *= d;
                }
                const Bsdf e = evaluate_opaque_bsdf(nx2, ny2, nz2, wox, woy, woz, pnx, pny, pnz, alpha2, metallic,
                                                    spec_prob, hit.ar, hit.ag, hit.ab);
                if (!(e.pdf > R(1e-12))) break;
                adj.scatter(hit, e, nx2, ny2, nz2, wox, woy, woz, pnx, pny, pnz, alpha2, metallic, accr, accg, accb);
                last_pdf = e.pdf;
                const R cos_i = nx2 * pnx + (ny2 * pny + nz2 * pnz);
                thr_r *= e.r * cos_i / e.pdf;
                thr_g *= e.g * cos_i / e.pdf;
                thr_b *= e.b * cos_i / e.pdf;
                prev_sumW = cur_sumW; prev_M = cur_M;
                prev_br = e.r; prev_bg = e.g; prev_bb = e.b; prev_cos = cos_i;
                ox = hit.p.x + pnx * R(0.001);
                oy = hit.p.y + pny * R(0.001);
                oz = hit.p.z + pnz * R(0.001);
                dx = nx2; dy = ny2; dz = nz2;
            }
            if (thr_r < 0) thr_r = 0;
            if (thr_g < 0) thr_g = 0;
            if (thr_b < 0) thr_b = 0;
            R mx = std::fmax(thr_r, thr_g);
            mx = std::fmax(mx, thr_b);
            if (!(mx >= R(1e-12))) break;
            if (bounce >= 1) {
                R rr = mx;
                if (rr < R(0.05)) rr = R(0.05);
                if (rr > R(0.9)) rr = R(0.9);
                if (!(urand() <= rr)) break;
                const R ir = R(1) / rr;
                thr_r *= ir; thr_g *= ir; thr_b *= ir;
            }
            bounce++;
        }
        return {accr, accg, accb};
    }

    // Camera ray direction. sample_sensor (pathtrace.metal:731) sums
    // sx^2 + (sy^2 + f^2); restir_generate (pathtrace.metal:834) sums
    // (sx^2 + sy^2) + f^2. Both groupings are kept.
    V3<R> primary_dir(R sx, R sy, const orc_camera& cam, bool generate_grouping = false) {
        const R focal = R(cam.focal);
        const R focal2 = focal * focal;
        const R dist = R(1) / std::sqrt(generate_grouping ? (sx * sx + sy * sy) + focal2 : sx * sx + (sy * sy + focal2));
        const R tx = sx * dist, ty = sy * dist, tz = focal * dist;
        return {tx * R(cam.m0) + (ty * R(cam.m1) + tz * R(cam.m2)), tx * R(cam.m3) + (ty * R(cam.m4) + tz * R(cam.m5)),
                tx * R(cam.m6) + (ty * R(cam.m7) + tz * R(cam.m8))};
    }
    V3<R> sample_sensor(R sx, R sy, const orc_camera& cam, int max_bounces) {
        return pathtrace({R(cam.x), R(cam.y), R(cam.z)}, primary_dir(sx, sy, cam), max_bounces);
    }
    void jitter(uint32_t x, uint32_t y, R& sx, R& sy) {  // pathtrace.metal:754-755
        const R res = R(P.resolution);
        sx = R(x) + R(1e-12) + (res - R(2e-12)) * urand() - R(P.width) * R(0.5);
        sy = R(P.height) * R(0.5) - (R(y) + R(1e-12) + (res - R(2e-12)) * urand());
    }
};

inline uint64_t pixel_seed(uint64_t seed, uint32_t x, uint32_t y) {
    return seed ^ (uint64_t(y + 1) * 0x9E3779B97F4A7C15ULL) ^ (uint64_t(x + 1) * 0xBF58476D1CE4E5B9ULL);
}
// The reference writes 0xD1B54A35A6F31B458UL (17 hex digits, pathtrace.metal:937);
// this build uses its low 64 bits.
constexpr uint64_t kSpatialSalt = 0x1B54A35A6F31B458ULL;

template <class R> inline uint8_t tonemap8(R c) {
    R v = std::floor(c / (c + 1) * 255);
    v = std::fmin(std::fmax(v, R(0)), R(255));
    return uint8_t(v);
}

// ---------------------------------------------------------------- kernels
struct Frame {
    const orc_scene* sc;
    orc_params P;
    uint64_t seed;
    Stats* stats;
};

template <class R> Tracer<R> make_tracer(const Frame& f, Stats* st) {
    Tracer<R> t;
    t.sc = f.sc;
    t.P = f.P;
    t.stats = st;
    return t;
}

// render_paths (pathtrace.metal:739-769); hdr receives sum/n before tonemapping.
template <class R> void render_paths(const Frame& f, uint32_t id, Stats* st, uint8_t* out, R* hdr) {
    const orc_params& P = f.P;
    const uint32_t x = id % uint32_t(P.width), y = id / uint32_t(P.width);
    Tracer<R> pt = make_tracer<R>(f, st);
    pt.rng_.seed(pixel_seed(f.seed, x, y));
    V3<R> sum;
    R sum_l = 0, sum_l2 = 0;
    int n = 0;
    for (int s = 0; s < P.spp; ++s) {
        R sx, sy;
        pt.jitter(x, y, sx, sy);
        V3<R> c = pt.sample_sensor(sx, sy, P.camera, P.bounces);
        if (P.clamp_value > 0) c = vmin(c, V3<R>(R(P.clamp_value), R(P.clamp_value), R(P.clamp_value)));
        sum = sum + c;
        ++n;
        if (P.adapt_min > 0) {
            const R lum = c.x * R(0.2126) + c.y * R(0.7152) + c.z * R(0.0722);
            sum_l += lum;
            sum_l2 += lum * lum;
            if (n >= P.adapt_min && n > 1 && n % P.adapt_step == 0) {
                const R mean = sum_l / R(n);
                const R var = std::fmax(R(0), (sum_l2 - sum_l * mean) / R(n - 1));
                const R threshold = R(P.adapt_rel) * mean + R(P.adapt_abs);
                if (var < threshold * threshold) break;
            }
        }
    }
    const V3<R> mean{sum.x / R(n), sum.y / R(n), sum.z / R(n)};
    out[3 * id] = tonemap8(mean.x); out[3 * id + 1] = tonemap8(mean.y); out[3 * id + 2] = tonemap8(mean.z);
    if (hdr) { hdr[3 * id] = mean.x; hdr[3 * id + 1] = mean.y; hdr[3 * id + 2] = mean.z; }
}

// render_paths_simd<L> (pathtrace.metal:773-815): L lanes per pixel, each with
// its own stream, combined by an xor butterfly after every block of L samples.
template <class R> void render_paths_lanes(const Frame& f, uint32_t pixel, int L, Stats* st, uint8_t* out, R* hdr) {
    const orc_params& P = f.P;
    const uint32_t x = pixel % uint32_t(P.width), y = pixel / uint32_t(P.width);
    std::vector<Tracer<R>> lanes(size_t(L), make_tracer<R>(f, st));
    for (int l = 0; l < L; ++l)
        lanes[size_t(l)].rng_.seed(pixel_seed(f.seed, x, y) ^ (uint64_t(l) * 0x94D049BB133111EBULL));
    V3<R> sum;
    R sum_l = 0, sum_l2 = 0;
    int n = 0;
    const size_t nl = size_t(L);
    std::vector<V3<R>> c(nl), c2(nl);
    std::vector<R> lum(nl), lum2(nl), t1(nl), t2(nl);
    for (int base = 0; base < P.spp; base += L) {
        for (int l = 0; l < L; ++l) {
            V3<R> v;
            if (base + l < P.spp) {
                R sx, sy;
                lanes[size_t(l)].jitter(x, y, sx, sy);
                v = lanes[size_t(l)].sample_sensor(sx, sy, P.camera, P.bounces);
                if (P.clamp_value > 0) v = vmin(v, V3<R>(R(P.clamp_value), R(P.clamp_value), R(P.clamp_value)));
            }
            c[size_t(l)] = v;
            lum[size_t(l)] = v.x * R(0.2126) + v.y * R(0.7152) + v.z * R(0.0722);
            lum2[size_t(l)] = lum[size_t(l)] * lum[size_t(l)];
        }
        for (int mask = 1; mask < L; mask *= 2) {
            for (int l = 0; l < L; ++l) {
                c2[size_t(l)] = c[size_t(l)] + c[size_t(l ^ mask)];
                t1[size_t(l)] = lum[size_t(l)] + lum[size_t(l ^ mask)];
                t2[size_t(l)] = lum2[size_t(l)] + lum2[size_t(l ^ mask)];
            }
            c.swap(c2); lum.swap(t1); lum2.swap(t2);
        }
        sum = sum + c[0];
        sum_l += lum[0];
        sum_l2 += lum2[0];
        n = std::min(base + L, P.spp);
        if (P.adapt_min > 0 && n >= P.adapt_min && n > 1 && n % P.adapt_step == 0) {
            const R mean = sum_l / R(n);
            const R var = std::fmax(R(0), (sum_l2 - sum_l * mean) / R(n - 1));
            const R threshold = R(P.adapt_rel) * mean + R(P.adapt_abs);
            if (var < threshold * threshold) break;
        }
    }
    const V3<R> mean{sum.x / R(n), sum.y / R(n), sum.z / R(n)};
    out[3 * pixel] = tonemap8(mean.x); out[3 * pixel + 1] = tonemap8(mean.y); out[3 * pixel + 2] = tonemap8(mean.z);
    if (hdr) { hdr[3 * pixel] = mean.x; hdr[3 * pixel + 1] = mean.y; hdr[3 * pixel + 2] = mean.z; }
}

// Path-replay adjoint of render_paths (not part of the reference): detached
// sampling (pdfs, MIS weights and roulette fixed), BSDF values and emission
// differentiated; each sample is replayed from its saved RNG state after a
// primal pass, and a bounce's scattering gets R * (df/dp) / f with R = C minus
// the radiance accumulated before it.
template <class R> struct AdjointHooks {
    const orc_scene* sc;
    const int32_t* tri_group;
    const float* base;  // 3 per group
    double* g;          // 5 per group
    V3<R> C, dC;
    static R drough(R r) { return (r > R(0.02) && r < R(1)) ? 4 * r * r * r : R(0); }
    V3<R> albedo_factor(const HitInfo<R>& hit) const {
        if (hit.shader == 1) return {0, 0, 0};
        const orc_triangle& t = sc->tris[hit.tri];
        return {t.ar != 0 ? hit.ar / R(t.ar) : R(0), t.ag != 0 ? hit.ag / R(t.ag) : R(0),
                t.ab != 0 ? hit.ab / R(t.ab) : R(0)};
    }
    void add(int grp, int k, R v) { g[grp * 5 + k] += double(v); }
    void emission(const HitInfo<R>& hit, R tr, R tg, R tb, R w) {
        const int grp = tri_group[hit.tri];
        if (grp < 0) return;
        const float* b = base + 3 * grp;
        add(grp, 4, (dC.x * tr * R(b[0]) + dC.y * tg * R(b[1]) + dC.z * tb * R(b[2])) * w);
    }
    template <class NO>
    void nee(const HitInfo<R>& hit, const NO& no, R tr, R tg, R tb, R wox, R woy, R woz, R nx, R ny, R nz, R alpha2,
             R metallic) {
        const int gl = tri_group[no.tri];
        if (gl >= 0) {
            const float* b = base + 3 * gl;
            add(gl, 4, (dC.x * tr * R(b[0]) * no.e[0] + dC.y * tg * R(b[1]) * no.e[1] + dC.z * tb * R(b[2]) * no.e[2]) *
                           no.k);
        }
        const int grp = tri_group[hit.tri];
        if (grp < 0) return;
        const auto d = Tracer<R>::evaluate_opaque_bsdf_grad(no.dir.x, no.dir.y, no.dir.z, wox, woy, woz, nx, ny, nz,
                                                            alpha2, metallic, hit.ar, hit.ag, hit.ab);
        const R w[3] = {dC.x * tr * no.le[0] * no.k, dC.y * tg * no.le[1] * no.k, dC.z * tb * no.le[2] * no.k};
        const V3<R> af = albedo_factor(hit);
        add(grp, 0, w[0] * d.da[0] * af.x);
        add(grp, 1, w[1] * d.da[1] * af.y);
        add(grp, 2, w[2] * d.da[2] * af.z);
        add(grp, 3, (w[0] * d.dl[0] + w[1] * d.dl[1] + w[2] * d.dl[2]) * drough(hit.rough));
    }
    template <class B>
    void scatter(const HitInfo<R>& hit, const B& e, R ix, R iy, R iz, R wox, R woy, R woz, R nx, R ny, R nz, R alpha2,
                 R metallic, R ar_, R ag_, R ab_) {
        const int grp = tri_group[hit.tri];
        if (grp < 0) return;
        const R rr = dC.x * (C.x - ar_), rg = dC.y * (C.y - ag_), rb = dC.z * (C.z - ab_);
        if (rr == 0 && rg == 0 && rb == 0) return;
        const auto d = Tracer<R>::evaluate_opaque_bsdf_grad(ix, iy, iz, wox, woy, woz, nx, ny, nz, alpha2, metallic,
                                                            hit.ar, hit.ag, hit.ab);
        const R q[3] = {e.r != 0 ? rr / e.r : R(0), e.g != 0 ? rg / e.g : R(0), e.b != 0 ? rb / e.b : R(0)};
        const V3<R> af = albedo_factor(hit);
        add(grp, 0, q[0] * d.da[0] * af.x);
        add(grp, 1, q[1] * d.da[1] * af.y);
        add(grp, 2, q[2] * d.da[2] * af.z);
        add(grp, 3, (q[0] * d.dl[0] + q[1] * d.dl[1] + q[2] * d.dl[2]) * drough(hit.rough));
    }
    template <class DS> void dielectric(const HitInfo<R>& hit, const DS& ds, R ar_, R ag_, R ab_) {
        if (ds.dlw_da2 == 0) return;
        const int grp = tri_group[hit.tri];
        if (grp < 0) return;
        add(grp, 3, (dC.x * (C.x - ar_) + dC.y * (C.y - ag_) + dC.z * (C.z - ab_)) * ds.dlw_da2 * drough(hit.rough));
    }
};

// The samples render_paths (lanes == 1) or render_paths_lanes drew for one
// pixel, replayed with the adjoint hooks; returns how many replays differed
// from their primal value.
template <class R>
uint64_t adjoint_pixel(const Frame& f, uint32_t pixel, int lanes, const float* d_hdr, const int32_t* tri_group,
                       const float* base, double* g) {
    const orc_params& P = f.P;
    const uint32_t x = pixel % uint32_t(P.width), y = pixel / uint32_t(P.width);
    const R fn = R(P.spp);
    const V3<R> dI{R(d_hdr[3 * pixel]) / fn, R(d_hdr[3 * pixel + 1]) / fn, R(d_hdr[3 * pixel + 2]) / fn};
    const V3<R> cam{R(P.camera.x), R(P.camera.y), R(P.camera.z)};
    uint64_t bad = 0;
    for (int l = 0; l < lanes; ++l) {
        Tracer<R> pt = make_tracer<R>(f, nullptr);
        uint64_t sd = pixel_seed(f.seed, x, y);
        if (lanes > 1) sd ^= uint64_t(l) * 0x94D049BB133111EBULL;
        pt.rng_.seed(sd);
        for (int k = l; k < P.spp; k += lanes) {
            const Rng st0 = pt.rng_;
            R sx, sy;
            pt.jitter(x, y, sx, sy);
            const V3<R> c = pt.sample_sensor(sx, sy, P.camera, P.bounces);
            V3<R> d = dI;
            if (P.clamp_value > 0) {
                if (c.x > R(P.clamp_value)) d.x = 0;
                if (c.y > R(P.clamp_value)) d.y = 0;
                if (c.z > R(P.clamp_value)) d.z = 0;
            }
            pt.rng_ = st0;
            pt.jitter(x, y, sx, sy);
            AdjointHooks<R> hk{f.sc, tri_group, base, g, c, d};
            const V3<R> c2 = pt.pathtrace_t(cam, pt.primary_dir(sx, sy, P.camera), P.bounces, hk);
            if (!(c2.x == c.x && c2.y == c.y && c2.z == c.z)) ++bad;
        }
    }
    return bad;
}

inline orc_vec3 fv(float x, float y, float z) { return {x, y, z}; }

void restir_generate(const Frame& f, uint32_t id, Stats* st, orc_gbuffer* gbuf, orc_reservoir* rinit) {
    const orc_params& P = f.P;
    const uint32_t x = id % uint32_t(P.width), y = id / uint32_t(P.width);
    Tracer<float> pt = make_tracer<float>(f, st);
    pt.rng_.seed(pixel_seed(f.seed, x, y));
    orc_gbuffer& g = gbuf[id];
    orc_reservoir& r0 = rinit[id];
    g.kind = 0;
    r0.valid = 0;
    float sx, sy;
    pt.jitter(x, y, sx, sy);
    const V3<float> rd = pt.primary_dir(sx, sy, P.camera, true);
    const V3<float> ro{P.camera.x, P.camera.y, P.camera.z};
    const HitInfo<float> hit = pt.cast_alpha_ray(ro, rd, 1e30f);
    if (hit.tri < 0) return;
    float pnx = hit.n_interp.x, pny = hit.n_interp.y, pnz = hit.n_interp.z;
    const float geom = hit.n_geom.x * pnx + (hit.n_geom.y * pny + hit.n_geom.z * pnz);
    if (geom < 0) { pnx = -pnx; pny = -pny; pnz = -pnz; }
    const float ndot = rd.x * pnx + (rd.y * pny + rd.z * pnz);
    if (ndot > 0) { pnx = -pnx; pny = -pny; pnz = -pnz; }
    const float wox = -rd.x, woy = -rd.y, woz = -rd.z;
    float rough = hit.rough;
    if (rough < 0) rough = 0;
    if (rough > 1) rough = 1;
    float metallic = f.sc->tris[hit.tri].metallic;
    if (metallic < 0) metallic = 0;
    if (metallic > 1) metallic = 1;
    g.pos = fv(hit.p.x, hit.p.y, hit.p.z);
    g.normal = fv(pnx, pny, pnz);
    g.albedo = fv(hit.ar, hit.ag, hit.ab);
    g.wo = fv(wox, woy, woz);
    g.emission = fv(hit.er, hit.eg, hit.eb);
    g.rough = rough;
    g.metallic = metallic;
    if (hit.ior > 1.0001f) {
        g.ior = hit.ior;
        g.kind = 3;
        return;
    }
    g.ior = 1;
    g.kind = 1;
    if (P.lights == 0 || !(P.light_total > 1e-12f)) return;
    int Mc = P.restir_candidates;
    if (Mc < 1) Mc = 1;
    if (Mc > 32) Mc = 32;
    float r2 = rough;
    if (r2 < 0.02f) r2 = 0.02f;
    const float alpha2 = (r2 * r2) * (r2 * r2);
    float spec_prob = 0.05f + 0.95f * metallic;
    if (spec_prob > 0.98f) spec_prob = 0.98f;
    if (metallic >= 0.999f) spec_prob = 1.0f;
    const float px = hit.p.x, py = hit.p.y, pz = hit.p.z;
    float sumWs = 0, sumWu = 0, sel_lx = 0, sel_ly = 0, sel_lz = 0, sel_nx = 0, sel_ny = 0, sel_nz = 0;
    float sel_er = 0, sel_eg = 0, sel_eb = 0, sel_alpha = 1, sel_ph = 0;
    bool have = false;
    const V3<float> so{px + pnx * 0.002f, py + pny * 0.002f, pz + pnz * 0.002f};
    for (int j = 0; j < Mc; ++j) {
        const float pick = pt.urand() * P.light_total;
        const orc_triangle& t = f.sc->tris[pt.pick_light(pick)];
        float lu = pt.urand(), lv = pt.urand();
        if (lu + lv > 1.0f) { lu = 1.0f - lu; lv = 1.0f - lv; }
        const float lw = 1.0f - lu - lv;
        const float qx = lw * t.v0.x + (lu * t.v1.x + lv * t.v2.x);
        const float qy = lw * t.v0.y + (lu * t.v1.y + lv * t.v2.y);
        const float qz = lw * t.v0.z + (lu * t.v1.z + lv * t.v2.z);
        const float ldx = qx - px, ldy = qy - py, ldz = qz - pz;
        const float dist = std::sqrt(ldx * ldx + ldy * ldy + ldz * ldz);
        float wu = 0, ws = 0, ph = 0;
        if (dist > 0.003f) {
            const float inv = 1.0f / dist;
            const float dxn = ldx * inv, dyn = ldy * inv, dzn = ldz * inv;
            const float lcos = -(t.nn.x * dxn + (t.nn.y * dyn + t.nn.z * dzn));
            const float scos = pnx * dxn + (pny * dyn + pnz * dzn);
            if (lcos > 1e-8f && scos > 1e-8f) {
                float ll = lum3<float>(t.er, t.eg, t.eb);
                if (ll < 1e-6f) ll = 1e-6f;
                const float q = dist * dist * ll / (lcos * P.light_total);
                if (q > 1e-12f) {
                    const auto e = pt.evaluate_opaque_bsdf(dxn, dyn, dzn, wox, woy, woz, pnx, pny, pnz, alpha2,
                                                           metallic, spec_prob, hit.ar, hit.ag, hit.ab);
                    const float Fr = t.er * e.r * scos * t.alpha, Fg = t.eg * e.g * scos * t.alpha,
                                Fb = t.eb * e.b * scos * t.alpha;
                    ph = lum3(Fr, Fg, Fb);
                    if (ph > 0) {
                        wu = ph / q;
                        if (!pt.occluded(so, {dxn, dyn, dzn}, dist - 0.003f)) ws = wu;
                    }
                }
            }
        }
        sumWu += wu;
        sumWs += ws;
        const float r = pt.urand();
        if (ws > 0 && r * sumWs < ws) {
            have = true;
            sel_lx = qx; sel_ly = qy; sel_lz = qz;
            sel_nx = t.nn.x; sel_ny = t.nn.y; sel_nz = t.nn.z;
            sel_er = t.er; sel_eg = t.eg; sel_eb = t.eb;
            sel_alpha = t.alpha;
            sel_ph = ph;
        }
    }
    if (!have || !(sumWs > 0) || !(sumWu > 0) || !(sel_ph > 0)) return;
    r0.light_pos = fv(sel_lx, sel_ly, sel_lz);
    r0.light_normal = fv(sel_nx, sel_ny, sel_nz);
    r0.emission = fv(sel_er, sel_eg, sel_eb);
    r0.alpha = sel_alpha;
    r0.p_hat = sel_ph;
    r0.sum_ws = sumWs;
    r0.sum_wu = sumWu;
    r0.m = uint32_t(Mc);
    r0.valid = 1;
}

// Reprojection of a world point into the previous camera (pathtrace.metal:957-966).
inline bool reproject(const orc_params& P, float px, float py, float pz, uint32_t& hidx) {
    const orc_camera& c = P.prev_camera;
    const float pdx = px - c.x, pdy = py - c.y, pdz = pz - c.z;
    const float ccx = c.m0 * pdx + c.m3 * pdy + c.m6 * pdz;
    const float ccy = c.m1 * pdx + c.m4 * pdy + c.m7 * pdz;
    const float ccz = c.m2 * pdx + c.m5 * pdy + c.m8 * pdz;
    if (!(ccz > 1e-6f)) return false;
    const float ppx = c.focal * ccx / ccz + float(P.width) * 0.5f;
    const float ppy = float(P.height) * 0.5f - c.focal * ccy / ccz;
    const int hx = int(std::floor(ppx)), hy = int(std::floor(ppy));
    if (!(hx >= 0 && hy >= 0 && hx < P.width && hy < P.height)) return false;
    hidx = uint32_t(hy) * uint32_t(P.width) + uint32_t(hx);
    return true;
}

void restir_spatial(const Frame& f, uint32_t id, Stats* st, const orc_gbuffer* gbuf, const orc_reservoir* rinit,
                    orc_reservoir* rout, const orc_gbuffer* hist_g, const orc_reservoir* hist_r) {
    const orc_params& P = f.P;
    const uint32_t x = id % uint32_t(P.width), y = id / uint32_t(P.width);
    rout[id] = rinit[id];
    if (gbuf[id].kind != 1 || rinit[id].valid == 0) return;
    int S = P.restir_spatial;
    if (S < 0) S = 0;
    if (S > 8) S = 8;
    if (S == 0) return;
    float radius = P.restir_radius;
    if (!(radius > 0)) radius = 16;
    int mcap = P.restir_mcap;
    if (mcap < 1) mcap = 1;
    Tracer<float> pt = make_tracer<float>(f, st);
    pt.rng_.seed(f.seed ^ (uint64_t(id + 1) * 0x9E3779B97F4A7C15ULL) ^ kSpatialSalt);
    const orc_gbuffer& g = gbuf[id];
    const V3<float> cpos = cvt<float>(g.pos), cnor = cvt<float>(g.normal), cwo = cvt<float>(g.wo),
                    calb = cvt<float>(g.albedo);
    float rough = g.rough;
    if (rough < 0) rough = 0;
    if (rough > 1) rough = 1;
    float metallic = g.metallic;
    if (metallic < 0) metallic = 0;
    if (metallic > 1) metallic = 1;
    float r2 = rough;
    if (r2 < 0.02f) r2 = 0.02f;
    const float alpha2 = (r2 * r2) * (r2 * r2);
    float spec_prob = 0.05f + 0.95f * metallic;
    if (spec_prob > 0.98f) spec_prob = 0.98f;
    if (metallic >= 0.999f) spec_prob = 1.0f;
    const orc_reservoir& ri = rinit[id];
    float cur_sumWs = ri.sum_ws, cur_sumWu = ri.sum_wu;
    uint32_t cur_M = ri.m;
    float cur_ph = ri.p_hat;
    orc_vec3 cur_lp = ri.light_pos, cur_ln = ri.light_normal, cur_le = ri.emission;
    float cur_alpha = ri.alpha;
    const V3<float> so{cpos.x + cnor.x * 0.002f, cpos.y + cnor.y * 0.002f, cpos.z + cnor.z * 0.002f};
    if (P.has_history != 0) {
        uint32_t hidx;
        if (reproject(P, cpos.x, cpos.y, cpos.z, hidx) && hist_g[hidx].kind == 1 && hist_r[hidx].valid != 0) {
            const V3<float> hp = cvt<float>(hist_g[hidx].pos), hn = cvt<float>(hist_g[hidx].normal);
            const float ddx = cpos.x - hp.x, ddy = cpos.y - hp.y, ddz = cpos.z - hp.z;
            const float dd = std::sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
            const float dcx = cpos.x - P.camera.x, dcy = cpos.y - P.camera.y, dcz = cpos.z - P.camera.z;
            const float dc = std::sqrt(dcx * dcx + dcy * dcy + dcz * dcz);
            const orc_reservoir& hr = hist_r[hidx];
            const float hsumWs = hr.sum_ws, hsumWu = hr.sum_wu, hph = hr.p_hat;
            const uint32_t hM = hr.m;
            if (dd < 0.02f * dc + 1e-3f && (cnor.x * hn.x + (cnor.y * hn.y + cnor.z * hn.z)) > 0.9f && hsumWu > 0 &&
                hM > 0 && hph > 0) {
                const V3<float> hlp = cvt<float>(hr.light_pos), hln = cvt<float>(hr.light_normal),
                                hle = cvt<float>(hr.emission);
                const float halpha = hr.alpha;
                const float ldx = hlp.x - cpos.x, ldy = hlp.y - cpos.y, ldz = hlp.z - cpos.z;
                const float dist = std::sqrt(ldx * ldx + ldy * ldy + ldz * ldz);
                if (dist > 0.003f) {
                    const float inv = 1.0f / dist;
                    const float dxn = ldx * inv, dyn = ldy * inv, dzn = ldz * inv;
                    const float lcos = -(hln.x * dxn + (hln.y * dyn + hln.z * dzn));
                    const float scos = cnor.x * dxn + (cnor.y * dyn + cnor.z * dzn);
                    if (lcos > 1e-8f && scos > 1e-8f) {
                        const auto e = pt.evaluate_opaque_bsdf(dxn, dyn, dzn, cwo.x, cwo.y, cwo.z, cnor.x, cnor.y,
                                                               cnor.z, alpha2, metallic, spec_prob, calb.x, calb.y,
                                                               calb.z);
                        const float ph = lum3(hle.x * e.r * scos * halpha, hle.y * e.g * scos * halpha,
                                              hle.z * e.b * scos * halpha);
                        if (ph > 0) {
                            const float wu = ph * hsumWu / hph;
                            const bool occ = pt.occluded(so, {dxn, dyn, dzn}, dist - 0.003f);
                            const float ws = occ ? 0.0f : ph * hsumWs / hph;
                            const float sum_s = cur_sumWs + ws, sum_u = cur_sumWu + wu;
                            uint32_t newM = cur_M + hM;
                            if (newM > uint32_t(mcap)) newM = uint32_t(mcap);
                            const float r = pt.urand();
                            if (sum_s > 0 && r * sum_s < ws) {
                                cur_lp = hr.light_pos; cur_ln = hr.light_normal; cur_le = hr.emission;
                                cur_alpha = halpha; cur_ph = ph;
                            }
                            cur_sumWs = sum_s; cur_sumWu = sum_u; cur_M = newM;
                        }
                    }
                }
            }
        }
    }
    for (int s = 0; s < S; ++s) {
        const float rx = pt.urand() * 2.0f - 1.0f, ry = pt.urand() * 2.0f - 1.0f;
        const int nx = int(float(x) + rx * radius + 0.5f), ny = int(float(y) + ry * radius + 0.5f);
        if (nx < 0 || ny < 0 || nx >= P.width || ny >= P.height) continue;
        if (nx == int(x) && ny == int(y)) continue;
        const uint32_t nidx = uint32_t(ny) * uint32_t(P.width) + uint32_t(nx);
        if (gbuf[nidx].kind != 1 || rinit[nidx].valid == 0) continue;
        const V3<float> nnor = cvt<float>(gbuf[nidx].normal);
        if (cnor.x * nnor.x + (cnor.y * nnor.y + cnor.z * nnor.z) < 0.9f) continue;
        const orc_reservoir& nr = rinit[nidx];
        const float nbr_sumWs = nr.sum_ws, nbr_sumWu = nr.sum_wu, nbr_ph = nr.p_hat;
        const uint32_t nbr_M = nr.m;
        if (!(nbr_sumWu > 0) || nbr_M == 0 || !(nbr_ph > 0)) continue;
        const V3<float> nlp = cvt<float>(nr.light_pos), nln = cvt<float>(nr.light_normal), nle = cvt<float>(nr.emission);
        const float nalpha = nr.alpha;
        const float ldx = nlp.x - cpos.x, ldy = nlp.y - cpos.y, ldz = nlp.z - cpos.z;
        const float dist = std::sqrt(ldx * ldx + ldy * ldy + ldz * ldz);
        if (!(dist > 0.003f)) continue;
        const float inv = 1.0f / dist;
        const float dxn = ldx * inv, dyn = ldy * inv, dzn = ldz * inv;
        const float lcos = -(nln.x * dxn + (nln.y * dyn + nln.z * dzn));
        const float scos = cnor.x * dxn + (cnor.y * dyn + cnor.z * dzn);
        if (!(lcos > 1e-8f && scos > 1e-8f)) continue;
        const auto e = pt.evaluate_opaque_bsdf(dxn, dyn, dzn, cwo.x, cwo.y, cwo.z, cnor.x, cnor.y, cnor.z, alpha2,
                                               metallic, spec_prob, calb.x, calb.y, calb.z);
        const float ph_cur = lum3(nle.x * e.r * scos * nalpha, nle.y * e.g * scos * nalpha, nle.z * e.b * scos * nalpha);
        if (!(ph_cur > 0)) continue;
        const float w_nbr_u = ph_cur * nbr_sumWu / nbr_ph;
        if (!(w_nbr_u > 0)) continue;
        const bool occ = pt.occluded(so, {dxn, dyn, dzn}, dist - 0.003f);
        const float w_nbr_s = occ ? 0.0f : ph_cur * nbr_sumWs / nbr_ph;
        const float sum_s = cur_sumWs + w_nbr_s, sum_u = cur_sumWu + w_nbr_u;
        if (!(sum_u > 0) || !(sum_s >= 0)) continue;
        const float r = pt.urand();
        uint32_t newM = cur_M + nbr_M;
        if (newM > uint32_t(mcap)) newM = uint32_t(mcap);
        if (sum_s > 0 && r * sum_s < w_nbr_s) {
            cur_lp = nr.light_pos; cur_ln = nr.light_normal; cur_le = nr.emission;
            cur_alpha = nalpha; cur_ph = ph_cur;
        }
        cur_sumWs = sum_s; cur_sumWu = sum_u; cur_M = newM;
    }
    orc_reservoir& o = rout[id];
    o.light_pos = cur_lp; o.light_normal = cur_ln; o.emission = cur_le; o.alpha = cur_alpha;
    o.p_hat = cur_ph; o.sum_ws = cur_sumWs; o.sum_wu = cur_sumWu; o.m = cur_M; o.valid = 1;
}

void restir_shade(const Frame& f, uint32_t id, Stats* st, const orc_gbuffer* gbuf, const orc_reservoir* rmerged,
                  float* hdr) {
    const orc_params& P = f.P;
    const uint32_t x = id % uint32_t(P.width), y = id / uint32_t(P.width);
    const orc_gbuffer& g = gbuf[id];
    if (g.kind == 0) { hdr[3 * id] = 0; hdr[3 * id + 1] = 0; hdr[3 * id + 2] = 0; return; }
    Tracer<float> pt = make_tracer<float>(f, st);
    pt.rng_.seed(f.seed ^ (uint64_t(id + 1) * 0xBF58476D1CE4E5B9ULL) ^ 0x94D049BB133111EBULL);
    const float cl = P.clamp_value;
    const V3<float> clampv{cl, cl, cl};
    if (g.kind == 3) {
        V3<float> sum;
        int n = 0;
        for (int s = 0; s < P.spp; ++s) {
            float sx, sy;
            pt.jitter(x, y, sx, sy);
            V3<float> c = pt.sample_sensor(sx, sy, P.camera, P.bounces);
            if (cl > 0) c = vmin(c, clampv);
            sum = sum + c;
            ++n;
        }
        const float fn = float(std::max(n, 1));
        V3<float> c{sum.x / fn, sum.y / fn, sum.z / fn};
        if (cl > 0) c = vmin(c, clampv);
        hdr[3 * id] = c.x; hdr[3 * id + 1] = c.y; hdr[3 * id + 2] = c.z;
        return;
    }
    const V3<float> gpos = cvt<float>(g.pos), gnor = cvt<float>(g.normal), galb = cvt<float>(g.albedo),
                    gwo = cvt<float>(g.wo), gemi = cvt<float>(g.emission);
    float rough = g.rough;
    if (rough < 0) rough = 0;
    if (rough > 1) rough = 1;
    float metallic = g.metallic;
    if (metallic < 0) metallic = 0;
    if (metallic > 1) metallic = 1;
    float r2 = rough;
    if (r2 < 0.02f) r2 = 0.02f;
    const float alpha2 = (r2 * r2) * (r2 * r2);
    float spec_prob = 0.05f + 0.95f * metallic;
    if (spec_prob > 0.98f) spec_prob = 0.98f;
    if (metallic >= 0.999f) spec_prob = 1.0f;
    const bool nomis = P.nomis != 0;
    V3<float> prim_direct;
    float m_sumWu = 0;
    uint32_t m_M = 0;
    const orc_reservoir& rm = rmerged[id];
    if (rm.valid != 0 && rm.m > 0 && rm.sum_wu > 0 && rm.sum_ws > 0 && rm.p_hat > 0) {
        const V3<float> lp = cvt<float>(rm.light_pos), ln = cvt<float>(rm.light_normal), le = cvt<float>(rm.emission);
        const float lalpha = rm.alpha;
        m_sumWu = rm.sum_wu;
        m_M = rm.m;
        const float m_sumWs = rm.sum_ws;
        const float ldx = lp.x - gpos.x, ldy = lp.y - gpos.y, ldz = lp.z - gpos.z;
        const float dist = std::sqrt(ldx * ldx + ldy * ldy + ldz * ldz);
        if (dist > 0.003f) {
            const float inv = 1.0f / dist;
            const float dxn = ldx * inv, dyn = ldy * inv, dzn = ldz * inv;
            const float lcos = -(ln.x * dxn + (ln.y * dyn + ln.z * dzn));
            const float scos = gnor.x * dxn + (gnor.y * dyn + gnor.z * dzn);
            if (lcos > 1e-8f && scos > 1e-8f) {
                const auto e = pt.evaluate_opaque_bsdf(dxn, dyn, dzn, gwo.x, gwo.y, gwo.z, gnor.x, gnor.y, gnor.z,
                                                       alpha2, metallic, spec_prob, galb.x, galb.y, galb.z);
                const float Fr = le.x * e.r * scos * lalpha, Fg = le.y * e.g * scos * lalpha,
                            Fb = le.z * e.b * scos * lalpha;
                const float ph = lum3(Fr, Fg, Fb);
                if (ph > 0) {
                    const float pdf_ris = float(m_M) * ph / m_sumWu;
                    float wnee = 1.0f;
                    if (!nomis && e.pdf > 1e-20f) wnee = pdf_ris / (pdf_ris + e.pdf);
                    const float W = m_sumWs / (float(m_M) * ph);
                    prim_direct = {Fr * W * wnee, Fg * W * wnee, Fb * W * wnee};
                }
            }
        }
    }
    if (P.bounces <= 1 || P.spp <= 0) {
        V3<float> c = gemi + prim_direct;
        if (cl > 0) c = vmin(c, clampv);
        hdr[3 * id] = c.x; hdr[3 * id + 1] = c.y; hdr[3 * id + 2] = c.z;
        return;
    }
    V3<float> sum;
    int n = 0;
    auto direct_only = [&]() {
        V3<float> full = gemi + prim_direct;
        if (cl > 0) full = vmin(full, clampv);
        sum = sum + full;
        ++n;
    };
    for (int s = 0; s < P.spp; ++s) {
        float nx2, ny2, nz2;
        int pdelta = 0;
        if (pt.urand() < spec_prob) {
            float hx, hy, hz;
            pt.sample_ggx_half(alpha2, gnor.x, gnor.y, gnor.z, hx, hy, hz);
            const float md = gwo.x * hx + (gwo.y * hy + gwo.z * hz);
            if (!(md > 1e-8f)) { direct_only(); continue; }
            nx2 = 2.0f * md * hx - gwo.x; ny2 = 2.0f * md * hy - gwo.y; nz2 = 2.0f * md * hz - gwo.z;
            pdelta = (rough <= 0.0201f) ? 1 : 0;
        } else {
            pt.sample_cosine_hemisphere(gnor.x, gnor.y, gnor.z, nx2, ny2, nz2);
            pdelta = 0;
        }
        const float ncos = nx2 * gnor.x + (ny2 * gnor.y + nz2 * gnor.z);
        if (!(ncos > 1e-8f)) { direct_only(); continue; }
        {
            const float d = 1.0f / std::sqrt(nx2 * nx2 + ny2 * ny2 + nz2 * nz2);
            nx2 *= d; ny2 *= d; nz2 *= d;
        }
        const auto e0 = pt.evaluate_opaque_bsdf(nx2, ny2, nz2, gwo.x, gwo.y, gwo.z, gnor.x, gnor.y, gnor.z, alpha2,
                                                metallic, spec_prob, galb.x, galb.y, galb.z);
        if (!(e0.pdf > 1e-12f)) { direct_only(); continue; }
        const float cos0 = nx2 * gnor.x + (ny2 * gnor.y + nz2 * gnor.z);
        V3<float> thr{e0.r * cos0 / e0.pdf, e0.g * cos0 / e0.pdf, e0.b * cos0 / e0.pdf};
        float ox = gpos.x + gnor.x * 0.001f, oy = gpos.y + gnor.y * 0.001f, oz = gpos.z + gnor.z * 0.001f;
        float dx = nx2, dy = ny2, dz = nz2;
        float last_pdf = e0.pdf;
        float pb_r = e0.r, pb_g = e0.g, pb_b = e0.b, pb_cos = cos0;
        float psumW = m_sumWu;
        uint32_t pM = m_M;
        int pdel = pdelta;
        V3<float> acc;
        for (int b = 1; b < P.bounces; ++b) {
            const HitInfo<float> hit = pt.cast_alpha_ray({ox, oy, oz}, {dx, dy, dz}, 1e30f);
            if (hit.tri < 0) break;
            float pnx = hit.n_interp.x, pny = hit.n_interp.y, pnz = hit.n_interp.z;
            const float gg = hit.n_geom.x * pnx + (hit.n_geom.y * pny + hit.n_geom.z * pnz);
            if (gg < 0) { pnx = -pnx; pny = -pny; pnz = -pnz; }
            const float nd = dx * pnx + (dy * pny + dz * pnz);
            bool front = true;
            if (nd > 0) { pnx = -pnx; pny = -pny; pnz = -pnz; front = false; }
            const float wox = -dx, woy = -dy, woz = -dz;
            float rh = hit.rough;
            if (rh < 0) rh = 0;
            if (rh > 1) rh = 1;
            if (rh < 0.02f) rh = 0.02f;
            const float al = rh * rh;
            const float a2 = al * al;
            float met = f.sc->tris[hit.tri].metallic;
            if (met < 0) met = 0;
            if (met > 1) met = 1;
            if (pdel == 1) {
                acc.x += hit.er * thr.x; acc.y += hit.eg * thr.y; acc.z += hit.eb * thr.z;
            } else {
                const float w = pt.emission_mis(hit, ox, oy, oz, last_pdf, psumW, pM, pb_r, pb_g, pb_b, pb_cos);
                acc.x += hit.er * thr.x * w; acc.y += hit.eg * thr.y * w; acc.z += hit.eb * thr.z * w;
            }
            if (hit.ior > 1.0001f) {
                const auto ds = pt.sample_dielectric(wox, woy, woz, pnx, pny, pnz, dx, dy, dz, hit.ior, front, rh, a2);
                if (!ds.valid) break;
                thr.x *= ds.eta_scale; thr.y *= ds.eta_scale; thr.z *= ds.eta_scale;
                float sgn = 1.0f;
                if (ds.n_x * pnx + (ds.n_y * pny + ds.n_z * pnz) < 0) sgn = -1.0f;
                ox = hit.p.x + pnx * 0.001f * sgn; oy = hit.p.y + pny * 0.001f * sgn; oz = hit.p.z + pnz * 0.001f * sgn;
                dx = ds.n_x; dy = ds.n_y; dz = ds.n_z;
                pdel = 1;
                psumW = 0; pM = 0;
            } else {
                float sp2 = 0.05f + 0.95f * met;
                if (sp2 > 0.98f) sp2 = 0.98f;
                if (met >= 0.999f) sp2 = 1.0f;
                float dlr = 0, dlg = 0, dlb = 0, csum = 0;
                uint32_t cM = 0;
                pt.direct_lighting(hit.p.x, hit.p.y, hit.p.z, pnx, pny, pnz, wox, woy, woz, rh, met, sp2, hit.ar,
                                   hit.ag, hit.ab, dlr, dlg, dlb, csum, cM);
                acc.x += dlr * thr.x; acc.y += dlg * thr.y; acc.z += dlb * thr.z;
                float qx2, qy2, qz2;
                if (pt.urand() < sp2) {
                    float hx, hy, hz;
                    pt.sample_ggx_half(a2, pnx, pny, pnz, hx, hy, hz);
                    const float md = wox * hx + (woy * hy + woz * hz);
                    if (!(md > 1e-8f)) break;
                    qx2 = 2.0f * md * hx - wox; qy2 = 2.0f * md * hy - woy; qz2 = 2.0f * md * hz - woz;
                    pdel = (rh <= 0.0201f) ? 1 : 0;
                } else {
                    pt.sample_cosine_hemisphere(pnx, pny, pnz, qx2, qy2, qz2);
                    pdel = 0;
                }
                const float nc = qx2 * pnx + (qy2 * pny + qz2 * pnz);
                if (!(nc > 1e-8f)) break;
                {
                    const float d = 1.0f / std::sqrt(qx2 * qx2 + qy2 * qy2 + qz2 * qz2);
                    qx2 *= d; qy2 *= d; qz2 *= d;
                }
                const auto ee = pt.evaluate_opaque_bsdf(qx2, qy2, qz2, wox, woy, woz, pnx, pny, pnz, a2, met, sp2,
                                                        hit.ar, hit.ag, hit.ab);
                if (!(ee.pdf > 1e-12f)) break;
                last_pdf = ee.pdf;
                const float ci = qx2 * pnx + (qy2 * pny + qz2 * pnz);
                thr.x *= ee.r * ci / ee.pdf; thr.y *= ee.g * ci / ee.pdf; thr.z *= ee.b * ci / ee.pdf;
                psumW = csum; pM = cM;
                pb_r = ee.r; pb_g = ee.g; pb_b = ee.b; pb_cos = ci;
                ox = hit.p.x + pnx * 0.001f; oy = hit.p.y + pny * 0.001f; oz = hit.p.z + pnz * 0.001f;
                dx = qx2; dy = qy2; dz = qz2;
            }
            float mx = std::fmax(thr.x, thr.y);
            mx = std::fmax(mx, thr.z);
            if (!(mx >= 1e-12f)) break;
            if (b >= 2) {
                float rr = mx;
                if (rr < 0.05f) rr = 0.05f;
                if (rr > 0.9f) rr = 0.9f;
                if (!(pt.urand() <= rr)) break;
                const float ir = 1.0f / rr;
                thr.x *= ir; thr.y *= ir; thr.z *= ir;
            }
        }
        V3<float> full = gemi + prim_direct + acc;
        if (cl > 0) full = vmin(full, clampv);
        sum = sum + full;
        ++n;
    }
    const float fn = float(std::max(n, 1));
    hdr[3 * id] = sum.x / fn; hdr[3 * id + 1] = sum.y / fn; hdr[3 * id + 2] = sum.z / fn;
}

inline float luma(V3<float> c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

void denoise_temporal(const orc_params& P, uint32_t id, const orc_gbuffer* gbuf, const orc_gbuffer* hist_g,
                      const float* hdr_in, const float* hist_col, float* den_out) {
    const uint32_t x = id % uint32_t(P.width), y = id / uint32_t(P.width);
    const V3<float> cur{hdr_in[3 * id], hdr_in[3 * id + 1], hdr_in[3 * id + 2]};
    V3<float> outc = cur;
    if (P.has_history != 0 && gbuf[id].kind == 1) {
        V3<float> mn = cur, mx = cur;
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int nx = int(x) + dx, ny = int(y) + dy;
                if (nx < 0 || ny < 0 || nx >= P.width || ny >= P.height) continue;
                const uint32_t n = uint32_t(ny) * uint32_t(P.width) + uint32_t(nx);
                const V3<float> v{hdr_in[3 * n], hdr_in[3 * n + 1], hdr_in[3 * n + 2]};
                mn = vmin(mn, v);
                mx = vmax(mx, v);
            }
        const V3<float> cpos = cvt<float>(gbuf[id].pos), cnor = cvt<float>(gbuf[id].normal);
        uint32_t hidx;
        if (reproject(P, cpos.x, cpos.y, cpos.z, hidx) && hist_g[hidx].kind == 1) {
            const V3<float> hp = cvt<float>(hist_g[hidx].pos), hn = cvt<float>(hist_g[hidx].normal);
            const float ddx = cpos.x - hp.x, ddy = cpos.y - hp.y, ddz = cpos.z - hp.z;
            const float dd = std::sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
            const float dcx = cpos.x - P.camera.x, dcy = cpos.y - P.camera.y, dcz = cpos.z - P.camera.z;
            const float dc = std::sqrt(dcx * dcx + dcy * dcy + dcz * dcz);
            if (dd < 0.02f * dc + 1e-3f && (cnor.x * hn.x + (cnor.y * hn.y + cnor.z * hn.z)) > 0.9f) {
                V3<float> h{hist_col[3 * hidx], hist_col[3 * hidx + 1], hist_col[3 * hidx + 2]};
                h = vmin(vmax(h, mn), mx);
                // Metal mix(h, cur, 0.2) = h + (cur - h) * 0.2
                outc = {h.x + (cur.x - h.x) * 0.2f, h.y + (cur.y - h.y) * 0.2f, h.z + (cur.z - h.z) * 0.2f};
            }
        }
    }
    den_out[3 * id] = outc.x; den_out[3 * id + 1] = outc.y; den_out[3 * id + 2] = outc.z;
}

void denoise_atrous(const orc_params& P, uint32_t id, int step, const orc_gbuffer* gbuf, const float* in, float* out) {
    const uint32_t x = id % uint32_t(P.width), y = id / uint32_t(P.width);
    if (gbuf[id].kind == 0) { out[3 * id] = 0; out[3 * id + 1] = 0; out[3 * id + 2] = 0; return; }
    const V3<float> c0{in[3 * id], in[3 * id + 1], in[3 * id + 2]};
    const V3<float> n0 = cvt<float>(gbuf[id].normal), p0 = cvt<float>(gbuf[id].pos), a0 = cvt<float>(gbuf[id].albedo);
    const float l0 = luma(c0);
    const float dcx = p0.x - P.camera.x, dcy = p0.y - P.camera.y, dcz = p0.z - P.camera.z;
    const float dc = std::sqrt(dcx * dcx + dcy * dcy + dcz * dcz);
    const int stride = 1 << step;
    V3<float> sum;
    float wsum = 0;
    for (int dy = -2; dy <= 2; ++dy)
        for (int dx = -2; dx <= 2; ++dx) {
            const int nx = int(x) + dx * stride, ny = int(y) + dy * stride;
            if (nx < 0 || ny < 0 || nx >= P.width || ny >= P.height) continue;
            const uint32_t n = uint32_t(ny) * uint32_t(P.width) + uint32_t(nx);
            if (gbuf[n].kind == 0) continue;
            const V3<float> c{in[3 * n], in[3 * n + 1], in[3 * n + 2]};
            const V3<float> nn = cvt<float>(gbuf[n].normal), pp = cvt<float>(gbuf[n].pos), aa = cvt<float>(gbuf[n].albedo);
            const float nd = nn.x * n0.x + (nn.y * n0.y + nn.z * n0.z);
            float wn = std::fmax(nd, 0.0f);
            wn = wn * wn; wn = wn * wn; wn = wn * wn; wn = wn * wn;  // ^16
            const float ex = pp.x - p0.x, ey = pp.y - p0.y, ez = pp.z - p0.z;
            // exp as rm_expf (one deterministic rounding, shared with the device)
            const float wz = rm_expf(-std::sqrt(ex * ex + ey * ey + ez * ez) / (0.1f * dc + 1e-4f));
            const float wl = rm_expf(-std::fabs(luma(c) - l0) / (0.3f * std::fmax(l0, 0.5f) + 1e-3f));
            const float ax = aa.x - a0.x, ay = aa.y - a0.y, az = aa.z - a0.z;
            const float wa = rm_expf(-std::sqrt(ax * ax + ay * ay + az * az) / 0.25f);
            const float w = wn * wz * wl * wa;
            sum = sum + c * w;
            wsum += w;
        }
    const V3<float> o = wsum > 1e-12f ? V3<float>{sum.x / wsum, sum.y / wsum, sum.z / wsum} : c0;
    out[3 * id] = o.x; out[3 * id + 1] = o.y; out[3 * id + 2] = o.z;
}

// ---------------------------------------------------------------- frame
void parallel_for(uint32_t n, int threads, const std::function<void(uint32_t)>& fn) {
    if (threads <= 1 || n < 64) {
        for (uint32_t i = 0; i < n; ++i) fn(i);
        return;
    }
    std::atomic<uint32_t> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&] {
            for (;;) {
                const uint32_t begin = next.fetch_add(64);
                if (begin >= n) break;
                const uint32_t end = std::min(n, begin + 64);
                for (uint32_t i = begin; i < end; ++i) fn(i);
            }
        });
    for (auto& t : pool) t.join();
}

struct History {
    std::vector<orc_gbuffer> gbuf;
    std::vector<orc_reservoir> res;
    std::vector<float> color;
    orc_camera cam{};
    int w = 0, h = 0;
    uint64_t seed = 0, hash = 0;
    bool valid = false;
};

uint64_t scene_hash(const orc_scene* sc) {  // metal.mm:241-248 over the uploaded records
    uint64_t h = uint64_t(sc->ntri) * 0x9E3779B97F4A7C15ULL ^ uint64_t(sc->nlight) * 0xBF58476D1CE4E5B9ULL;
    uint32_t lt;
    std::memcpy(&lt, &sc->light_total, 4);
    h ^= uint64_t(lt);
    const unsigned char* b = reinterpret_cast<const unsigned char*>(sc->tris);
    for (size_t i = 0, n = sc->ntri * sizeof(orc_triangle); i < n; ++i) {
        h ^= uint64_t(b[i]);
        h *= 0x100000001B3ULL;
    }
    return h;
}

}  // namespace orc

extern "C" {

typedef struct {
    orc_gbuffer* gbuf;       // W*H, may be NULL
    orc_reservoir* rinit;    // W*H, may be NULL
    orc_reservoir* rmerged;  // W*H, may be NULL
    float* shade;            // W*H*3 restir_shade output, may be NULL
    int32_t* path;           // out: 0 megakernel, 1 restir
    int32_t* lanes;          // out: megakernel lanes
    int32_t* has_history;    // out
    uint64_t* stats;         // out: [rays, node visits, triangle tests, then rays of the megakernel,
                             // generate, spatial and shade passes], 7 words, may be NULL
} orc_debug;

void* orc_history_create(void) { return new orc::History(); }
void orc_history_destroy(void* h) { delete static_cast<orc::History*>(h); }

// One frame with the semantics of render_image_metal (metal.mm:95-414).
// Returns 0, or -1 for invalid dimensions. rgb8: W*H*3; hdr (optional): W*H*3
// linear values that were tonemapped.
int orc_render_frame(const orc_scene* sc, const orc_frame* f, uint64_t seed, void* hist_handle, int threads,
                     uint8_t* rgb8, float* hdr, orc_debug* dbg) {
    using namespace orc;
    if (f->width <= 0 || f->height <= 0 || f->spp <= 0 || f->bounces < 0 ||
        size_t(f->width) * size_t(f->height) > 100000000)
        return -1;
    History* hist = static_cast<History*>(hist_handle);
    History scratch_hist;
    if (!hist) hist = &scratch_hist;
    const uint32_t px = uint32_t(f->width) * uint32_t(f->height);
    orc_params P{};
    P.camera = f->camera;
    P.width = f->width; P.height = f->height; P.spp = f->spp; P.bounces = f->bounces;
    P.nodes = int32_t(sc->nnode); P.lights = int32_t(sc->nlight);
    P.adapt_min = f->adapt_min;
    P.adapt_step = f->adapt_step > 0 ? f->adapt_step : 8;
    P.nomis = f->nomis;
    P.opaque = 1;
    for (size_t i = 0; i < sc->ntri; ++i)
        if (!(sc->tris[i].alpha >= 1)) { P.opaque = 0; break; }
    P.light_total = sc->light_total;
    P.resolution = f->resolution; P.clamp_value = f->clamp_value;
    P.adapt_rel = f->adapt_rel; P.adapt_abs = f->adapt_abs;
    P.restir_candidates = f->restir_candidates; P.restir_spatial = f->restir_spatial;
    P.restir_mcap = f->restir_mcap; P.restir_radius = f->restir_radius;
    const size_t restir_bytes = size_t(px) * (sizeof(orc_gbuffer) + 2 * sizeof(orc_reservoir));
    const bool use_spatial = P.restir_spatial > 0 && P.lights > 0 && P.bounces > 0 && P.opaque &&
                             restir_bytes <= size_t(512) * 1024 * 1024;
    const uint64_t hash = scene_hash(sc);
    bool have_history = use_spatial && hist->valid && hist->w == f->width && hist->h == f->height &&
                        seed != hist->seed && hash == hist->hash;
    if (!f->temporal) have_history = false;
    P.has_history = have_history ? 1 : 0;
    P.prev_camera = hist->cam;
    Stats total;
    std::vector<Stats> per(px);
    Frame fr{sc, P, seed, nullptr};
    const bool want_stats = dbg && dbg->stats;
    if (dbg && dbg->has_history) *dbg->has_history = P.has_history;
    if (use_spatial) {
        if (dbg && dbg->path) *dbg->path = 1;
        std::vector<orc_gbuffer> gbuf(px);
        std::vector<orc_reservoir> rinit(px), rmerged(px);
        std::vector<float> shade(size_t(px) * 3);
        if (hist->gbuf.size() != px) { hist->gbuf.assign(px, orc_gbuffer{}); hist->res.assign(px, orc_reservoir{}); }
        if (hist->color.size() != size_t(px) * 3) hist->color.assign(size_t(px) * 3, 0.0f);
        auto rays_so_far = [&] {
            uint64_t r = 0;
            for (const auto& s : per) r += s.rays;
            return r;
        };
        uint64_t pass_rays[3] = {0, 0, 0};
        parallel_for(px, threads, [&](uint32_t i) { restir_generate(fr, i, want_stats ? &per[i] : nullptr, gbuf.data(), rinit.data()); });
        if (want_stats) pass_rays[0] = rays_so_far();
        parallel_for(px, threads, [&](uint32_t i) {
            restir_spatial(fr, i, want_stats ? &per[i] : nullptr, gbuf.data(), rinit.data(), rmerged.data(),
                           hist->gbuf.data(), hist->res.data());
        });
        if (want_stats) pass_rays[1] = rays_so_far() - pass_rays[0];
        parallel_for(px, threads, [&](uint32_t i) { restir_shade(fr, i, want_stats ? &per[i] : nullptr, gbuf.data(), rmerged.data(), shade.data()); });
        if (want_stats) {
            pass_rays[2] = rays_so_far() - pass_rays[0] - pass_rays[1];
            dbg->stats[3] = 0;
            dbg->stats[4] = pass_rays[0]; dbg->stats[5] = pass_rays[1]; dbg->stats[6] = pass_rays[2];
        }
        std::vector<float> final_img;
        if (f->denoise) {
            std::vector<float> a(size_t(px) * 3), b(size_t(px) * 3);
            parallel_for(px, threads, [&](uint32_t i) {
                denoise_temporal(P, i, gbuf.data(), hist->gbuf.data(), shade.data(), hist->color.data(), a.data());
            });
            const int iters = std::max(0, f->denoise_iterations);
            for (int s = 0; s < iters; ++s) {
                parallel_for(px, threads, [&](uint32_t i) { denoise_atrous(P, i, s, gbuf.data(), a.data(), b.data()); });
                a.swap(b);
            }
            final_img = a;
            hist->color = final_img;
        } else {
            final_img = shade;
        }
        for (uint32_t i = 0; i < px; ++i)
            for (int c = 0; c < 3; ++c) rgb8[3 * i + c] = tonemap8(final_img[3 * size_t(i) + c]);
        if (hdr) std::memcpy(hdr, final_img.data(), final_img.size() * sizeof(float));
        if (dbg && dbg->gbuf) std::memcpy(dbg->gbuf, gbuf.data(), px * sizeof(orc_gbuffer));
        if (dbg && dbg->rinit) std::memcpy(dbg->rinit, rinit.data(), px * sizeof(orc_reservoir));
        if (dbg && dbg->rmerged) std::memcpy(dbg->rmerged, rmerged.data(), px * sizeof(orc_reservoir));
        if (dbg && dbg->shade) std::memcpy(dbg->shade, shade.data(), shade.size() * sizeof(float));
        hist->gbuf.swap(gbuf);
        hist->res.swap(rmerged);
        hist->cam = f->camera;
        hist->w = f->width; hist->h = f->height;
        hist->seed = seed; hist->hash = hash;
        hist->valid = true;
    } else {
        int lanes = 1;
        if (f->spp >= 8 && (P.adapt_min <= 0 || P.adapt_step % 8 == 0)) lanes = 8;
        if (f->spp >= 32 && (P.adapt_min <= 0 || P.adapt_step % 32 == 0)) lanes = 32;
        if (dbg && dbg->path) *dbg->path = 0;
        if (dbg && dbg->lanes) *dbg->lanes = lanes;
        std::vector<float> img(size_t(px) * 3);
        parallel_for(px, threads, [&](uint32_t i) {
            if (lanes == 1) render_paths<float>(fr, i, want_stats ? &per[i] : nullptr, rgb8, img.data());
            else render_paths_lanes<float>(fr, i, lanes, want_stats ? &per[i] : nullptr, rgb8, img.data());
        });
        if (hdr) std::memcpy(hdr, img.data(), img.size() * sizeof(float));
    }
    if (want_stats) {
        for (const auto& s : per) { total.rays += s.rays; total.nodes += s.nodes; total.tris += s.tris; }
        dbg->stats[0] = total.rays; dbg->stats[1] = total.nodes; dbg->stats[2] = total.tris;
        if (!use_spatial) { dbg->stats[3] = total.rays; dbg->stats[4] = dbg->stats[5] = dbg->stats[6] = 0; }
    }
    return 0;
}

// restir_spatial alone (pathtrace.metal:919-1054) over caller-supplied
// G-buffer and initial reservoirs (W*H each); temporal reuse from hist_g /
// hist_r through prev_camera when all three are given. Shadow rays are traced
// against the scene. rout receives W*H merged reservoirs. Returns 0 or -1.
int orc_restir_merge(const orc_scene* sc, const orc_frame* f, uint64_t seed, const orc_gbuffer* gbuf,
                     const orc_reservoir* rinit, const orc_gbuffer* hist_g, const orc_reservoir* hist_r,
                     const orc_camera* prev_camera, int threads, orc_reservoir* rout) {
    using namespace orc;
    if (f->width <= 0 || f->height <= 0 || !gbuf || !rinit || !rout) return -1;
    const bool hist = hist_g && hist_r && prev_camera;
    orc_params P{};
    P.camera = f->camera;
    P.width = f->width; P.height = f->height; P.spp = f->spp; P.bounces = f->bounces;
    P.nodes = int32_t(sc->nnode); P.lights = int32_t(sc->nlight);
    P.nomis = f->nomis;
    P.opaque = 1;
    for (size_t i = 0; i < sc->ntri; ++i)
        if (!(sc->tris[i].alpha >= 1)) { P.opaque = 0; break; }
    P.light_total = sc->light_total;
    P.resolution = f->resolution; P.clamp_value = f->clamp_value;
    P.restir_candidates = f->restir_candidates; P.restir_spatial = f->restir_spatial;
    P.restir_mcap = f->restir_mcap; P.restir_radius = f->restir_radius;
    P.has_history = hist ? 1 : 0;
    if (hist) P.prev_camera = *prev_camera;
    Frame fr{sc, P, seed, nullptr};
    const uint32_t px = uint32_t(f->width) * uint32_t(f->height);
    parallel_for(px, threads, [&](uint32_t i) {
        restir_spatial(fr, i, nullptr, gbuf, rinit, rout, hist ? hist_g : gbuf, hist ? hist_r : rinit);
    });
    return 0;
}

// The denoiser alone: denoise_temporal, then f->denoise_iterations a-trous
// passes (pathtrace.metal:1263-1354) over a linear image (W*H*3) and its
// G-buffer; the temporal pass blends hist_color (seen through prev_camera,
// with hist_g) when all three are given. out: W*H*3. Returns 0 or -1.
int orc_denoise_image(const orc_frame* f, const float* color, const orc_gbuffer* gbuf, const float* hist_color,
                      const orc_gbuffer* hist_g, const orc_camera* prev_camera, int threads, float* out) {
    using namespace orc;
    if (f->width <= 0 || f->height <= 0 || !color || !gbuf || !out) return -1;
    const bool hist = hist_color && hist_g && prev_camera;
    orc_params P{};
    P.camera = f->camera;
    P.width = f->width; P.height = f->height;
    P.has_history = hist ? 1 : 0;
    if (hist) P.prev_camera = *prev_camera;
    const uint32_t px = uint32_t(f->width) * uint32_t(f->height);
    std::vector<float> a(size_t(px) * 3), b(size_t(px) * 3);
    parallel_for(px, threads, [&](uint32_t i) {
        denoise_temporal(P, i, gbuf, hist ? hist_g : gbuf, color, hist ? hist_color : color, a.data());
    });
    for (int s = 0; s < std::max(0, f->denoise_iterations); ++s) {
        parallel_for(px, threads, [&](uint32_t i) { denoise_atrous(P, i, s, gbuf, a.data(), b.data()); });
        a.swap(b);
    }
    std::memcpy(out, a.data(), a.size() * sizeof(float));
    return 0;
}

// Plain megakernel render (render_paths semantics, no adaptivity, lanes=1) in
// FP64 with FP64 accumulation; hdr receives the per-pixel mean (W*H*3).
int orc_render_reference_f64(const orc_scene* sc, const orc_frame* f, uint64_t seed, int threads, double* hdr) {
    using namespace orc;
    orc_params P{};
    P.camera = f->camera;
    P.width = f->width; P.height = f->height; P.spp = f->spp; P.bounces = f->bounces;
    P.nodes = int32_t(sc->nnode); P.lights = int32_t(sc->nlight);
    P.adapt_min = 0; P.adapt_step = 8; P.nomis = f->nomis;
    P.opaque = 1;
    for (size_t i = 0; i < sc->ntri; ++i)
        if (!(sc->tris[i].alpha >= 1)) { P.opaque = 0; break; }
    P.light_total = sc->light_total;
    P.resolution = f->resolution; P.clamp_value = f->clamp_value;
    P.restir_candidates = f->restir_candidates; P.restir_spatial = 0;
    P.restir_mcap = f->restir_mcap; P.restir_radius = f->restir_radius;
    Frame fr{sc, P, seed, nullptr};
    const uint32_t px = uint32_t(f->width) * uint32_t(f->height);
    std::vector<uint8_t> dummy(size_t(px) * 3);
    parallel_for(px, threads, [&](uint32_t i) { render_paths<double>(fr, i, nullptr, dummy.data(), hdr); });
    return 0;
}

// Adjoint of the path-tracing frame orc_render_frame draws for (f, seed): grad
// (5 doubles per group: albedo r, g, b, roughness, emission scale) of
// sum d_hdr . hdr. precision 32 follows the HIP kernel's arithmetic; 64
// evaluates the same estimator in double. Returns -1 for a frame the adjoint
// does not cover, else the number of samples whose replay differed.
long long orc_render_adjoint(const orc_scene* sc, const orc_frame* f, uint64_t seed, const float* d_hdr,
                             const int32_t* tri_group, const float* group_emission, int ngroup, int precision,
                             int threads, double* grad) {
    using namespace orc;
    if (f->width <= 0 || f->height <= 0 || f->spp <= 0 || f->bounces < 0 || ngroup < 1) return -1;
    const uint32_t px = uint32_t(f->width) * uint32_t(f->height);
    orc_params P{};
    P.camera = f->camera;
    P.width = f->width; P.height = f->height; P.spp = f->spp; P.bounces = f->bounces;
    P.nodes = int32_t(sc->nnode); P.lights = int32_t(sc->nlight);
    P.adapt_min = f->adapt_min;
    P.adapt_step = f->adapt_step > 0 ? f->adapt_step : 8;
    P.nomis = f->nomis;
    P.opaque = 1;
    for (size_t i = 0; i < sc->ntri; ++i)
        if (!(sc->tris[i].alpha >= 1)) { P.opaque = 0; break; }
    P.light_total = sc->light_total;
    P.resolution = f->resolution; P.clamp_value = f->clamp_value;
    P.adapt_rel = f->adapt_rel; P.adapt_abs = f->adapt_abs;
    P.restir_candidates = f->restir_candidates; P.restir_spatial = f->restir_spatial;
    P.restir_mcap = f->restir_mcap; P.restir_radius = f->restir_radius;
    const size_t restir_bytes = size_t(px) * (sizeof(orc_gbuffer) + 2 * sizeof(orc_reservoir));
    const bool use_spatial = P.restir_spatial > 0 && P.lights > 0 && P.bounces > 0 && P.opaque &&
                             restir_bytes <= size_t(512) * 1024 * 1024;
    if (use_spatial || P.restir_candidates > 1 || (P.adapt_min > 0 && P.adapt_min <= P.spp)) return -1;
    int lanes = 1;
    if (f->spp >= 8 && (P.adapt_min <= 0 || P.adapt_step % 8 == 0)) lanes = 8;
    if (f->spp >= 32 && (P.adapt_min <= 0 || P.adapt_step % 32 == 0)) lanes = 32;
    Frame fr{sc, P, seed, nullptr};
    std::mutex mu;
    std::vector<double> total(size_t(ngroup) * 5, 0.0);
    uint64_t bad = 0;
    parallel_for(px, threads, [&](uint32_t i) {
        std::vector<double> g(size_t(ngroup) * 5, 0.0);
        const uint64_t b = precision == 64
                               ? adjoint_pixel<double>(fr, i, lanes, d_hdr, tri_group, group_emission, g.data())
                               : adjoint_pixel<float>(fr, i, lanes, d_hdr, tri_group, group_emission, g.data());
        std::lock_guard<std::mutex> lock(mu);
        for (size_t k = 0; k < g.size(); ++k) total[k] += g[k];
        bad += b;
    });
    std::memcpy(grad, total.data(), total.size() * sizeof(double));
    return (long long)bad;
}

// Closest-hit queries: tri/t/u/v per ray (tri = -1 on a miss). any != 0 runs
// shadow (any-hit) queries instead: tri >= 0 means occluded. stats: [nodes, tris].
int orc_trace(const orc_scene* sc, const float* o, const float* d, const float* tmax, size_t n, int any,
              int threads, int32_t* tri, float* t, float* u, float* v, uint64_t* stats) {
    using namespace orc;
    orc_params P{};
    P.nodes = int32_t(sc->nnode);
    std::vector<Stats> per(stats ? n : 0);
    parallel_for(uint32_t(n), threads, [&](uint32_t i) {
        Tracer<float> tr;
        tr.sc = sc;
        tr.P = P;
        tr.stats = stats ? &per[i] : nullptr;
        const V3<float> ro{o[3 * i], o[3 * i + 1], o[3 * i + 2]}, rd{d[3 * i], d[3 * i + 1], d[3 * i + 2]};
        const HitInfo<float> h = tr.intersect(ro, rd, tmax ? tmax[i] : 1e30f, any != 0);
        tri[i] = h.tri;
        if (t) t[i] = h.t;
        if (u) u[i] = h.u;
        if (v) v[i] = h.v;
    });
    if (stats) {
        stats[0] = stats[1] = 0;
        for (const auto& s : per) { stats[0] += s.nodes; stats[1] += s.tris; }
    }
    return 0;
}

// Host walk of the backend's own traversal tree, to count what its traversal
// fetches. nodes: nnode 4-wide nodes of 8 float4 words each (lower bounds x,
// upper x, lower y, upper y, lower z, upper z of the four children, then
// their references and kinds as int bits: -1 empty, 0 inner node, k > 0 leaf
// of k triangles from leaf-order index ref); geo: the triangles in leaf order,
// 3 float4 words each (v0.xyz v1.x | v1.yz v2.xy | v2.z id). The walk is the
// device's (Tracer::intersect): slab tests with fused multiply-adds, children
// sorted by entry distance, hit leaves intersected nearest-first within the
// current hit plus 2^-12, inner children descended nearest-first with the rest
// stacked. stats: [node visits, triangle tests] over all rays.
int orc_trace_wide(const float* nodes, size_t nnode, const float* geo, const float* o, const float* d,
                   const float* tmax, size_t n, int any, int threads, int32_t* tri, float* t, uint64_t* stats) {
    using namespace orc;
    std::vector<Stats> per(n);
    const WideTree w{nodes, nnode, geo};
    parallel_for(uint32_t(n), threads, [&](uint32_t i) {
        const V3<float> ro{o[3 * i], o[3 * i + 1], o[3 * i + 2]}, rd{d[3 * i], d[3 * i + 1], d[3 * i + 2]};
        Stats& st = per[i];
        ++st.rays;
        float ht, u, v;
        tri[i] = wide_walk(w, ro, rd, tmax ? tmax[i] : 1e30f, any != 0, st, ht, u, v);
        if (t) t[i] = ht;
    });
    stats[0] = stats[1] = 0;
    for (const auto& s : per) { stats[0] += s.nodes; stats[1] += s.tris; }
    return 0;
}

// FP32 frames (orc_render_frame, orc_restir_merge, orc_render_adjoint) trace
// the given 4-wide tree with the backend's traversal until this is called
// again with nodes = NULL (process-wide; not for concurrent callers).
void orc_set_wide_tree(const float* nodes, size_t nnode, const float* geo) {
    orc::g_wide = orc::WideTree{nodes, nodes ? nnode : 0, nodes ? geo : nullptr};
}

// Brute force over every triangle with the same Moller-Trumbore arithmetic and
// tie-break (closest distance, then smallest index).
int orc_trace_brute(const orc_scene* sc, const float* o, const float* d, size_t n, int threads, int32_t* tri,
                    float* t) {
    using namespace orc;
    parallel_for(uint32_t(n), threads, [&](uint32_t i) {
        const V3<float> ro{o[3 * i], o[3 * i + 1], o[3 * i + 2]}, rd{d[3 * i], d[3 * i + 1], d[3 * i + 2]};
        float best = 1e30f;
        int best_j = -1;
        for (size_t j = 0; j < sc->ntri; ++j) {
            const orc_triangle& tr = sc->tris[j];
            const V3<float> v0 = cvt<float>(tr.v0);
            const V3<float> e1 = cvt<float>(tr.v1) - v0, e2 = cvt<float>(tr.v2) - v0;
            const V3<float> p = cross3(rd, e2);
            const float det = dot3(e1, p);
            if (std::fabs(det) < 1e-12f) continue;
            const V3<float> tv = ro - v0;
            const float uu = dot3(tv, p) / det;
            if (uu < 0 || uu > 1) continue;
            const V3<float> q = cross3(tv, e1);
            const float vv = dot3(rd, q) / det;
            if (vv < 0 || uu + vv > 1) continue;
            const float dd = dot3(e2, q) / det;
            if (dd > 1e-8f && (dd < best || (dd == best && best_j >= 0 && int(j) < best_j))) {
                best = dd;
                best_j = int(j);
            }
        }
        tri[i] = best_j;
        if (t) t[i] = best;
    });
    return 0;
}

// RNG known-answer helper: n draws from a stream seeded with s.
void orc_rng_draws(uint64_t s, int n, float* out, uint32_t* states) {
    orc::Rng r;
    r.seed(s);
    if (states) states[0] = r.state;
    for (int i = 0; i < n; ++i) {
        out[i] = r.uniform01();
        if (states) states[i + 1] = r.state;
    }
}

void orc_sincos(const float* x, size_t n, float* s, float* c) {
    for (size_t i = 0; i < n; ++i) rm_sincosf(x[i], &s[i], &c[i]);
}

// Every float whose bit pattern with the sign cleared lies in [lo, hi):
// the number of them (both signs) for which rm_div_pi differs from the
// IEEE quotient x / pi (a pair of NaNs counts as equal).
uint64_t orc_check_div_pi(uint32_t lo, uint32_t hi, int threads) {
    const float kpi = 3.14159265358979323846f;
    std::atomic<uint64_t> bad{0};
    const uint32_t n = hi > lo ? hi - lo : 0, chunk = 1u << 20;
    orc::parallel_for((n + chunk - 1) / chunk, threads, [&](uint32_t c) {
        uint64_t local = 0;
        const uint32_t a = lo + c * chunk, b = std::min(hi, a + chunk);
        for (uint32_t u = a; u < b; ++u)
            for (uint32_t sign = 0; sign < 2; ++sign) {
                const uint32_t bits = u | (sign << 31);
                float x;
                std::memcpy(&x, &bits, 4);
                const float want = x / kpi, got = rm_div_pi(x);
                uint32_t wb, gb;
                std::memcpy(&wb, &want, 4);
                std::memcpy(&gb, &got, 4);
                if (wb != gb && !(want != want && got != got)) ++local;
            }
        bad += local;
    });
    return bad.load();
}

// rm_expf against libm's double exp rounded to float, over the float bit
// patterns [lo, hi) of both signs: returns the largest difference in ulp
// (2^31 where one result is NaN and the other is not), and stores how many
// results differ at all.
uint64_t orc_check_exp(uint32_t lo, uint32_t hi, int threads, uint64_t* differ) {
    std::atomic<uint64_t> worst{0}, ndiff{0};
    const uint32_t n = hi > lo ? hi - lo : 0, chunk = 1u << 20;
    orc::parallel_for((n + chunk - 1) / chunk, threads, [&](uint32_t c) {
        uint64_t lw = 0, ld = 0;
        const uint32_t a = lo + c * chunk, b = std::min(hi, a + chunk);
        for (uint32_t u = a; u < b; ++u)
            for (uint32_t sign = 0; sign < 2; ++sign) {
                const uint32_t bits = u | (sign << 31);
                float x;
                std::memcpy(&x, &bits, 4);
                const float want = float(std::exp(double(x))), got = rm_expf(x);
                if (want != want || got != got) {
                    if ((want != want) != (got != got)) lw = uint64_t(1) << 31, ++ld;
                    continue;
                }
                int32_t wb, gb;  // both results are >= 0: bit patterns order like values
                std::memcpy(&wb, &want, 4);
                std::memcpy(&gb, &got, 4);
                const uint64_t d = uint64_t(wb > gb ? wb - gb : gb - wb);
                lw = std::max(lw, d);
                ld += d != 0;
            }
        uint64_t cur = worst.load();
        while (lw > cur && !worst.compare_exchange_weak(cur, lw)) {}
        ndiff += ld;
    });
    if (differ) *differ = ndiff.load();
    return worst.load();
}

// rm_atanf against libm's double atan rounded to float, over the float bit
// patterns [lo, hi) of both signs: the largest difference in ulp (2^31 where
// one result is NaN and the other is not); stores how many results differ.
uint64_t orc_check_atan(uint32_t lo, uint32_t hi, int threads, uint64_t* differ) {
    std::atomic<uint64_t> worst{0}, ndiff{0};
    const uint32_t n = hi > lo ? hi - lo : 0, chunk = 1u << 20;
    orc::parallel_for((n + chunk - 1) / chunk, threads, [&](uint32_t c) {
        uint64_t lw = 0, ld = 0;
        const uint32_t a = lo + c * chunk, b = std::min(hi, a + chunk);
        for (uint32_t u = a; u < b; ++u) {
            float x;
            std::memcpy(&x, &u, 4);  // x >= 0; atan is odd and rm_atanf mirrors the sign exactly
            const float want = float(std::atan(double(x))), got = rm_atanf(x), gneg = rm_atanf(-x);
            if (want != want || got != got) {
                if ((want != want) != (got != got)) lw = uint64_t(1) << 31, ++ld;
                continue;
            }
            if (!(gneg == -got)) lw = uint64_t(1) << 31, ++ld;
            int32_t wb, gb;
            std::memcpy(&wb, &want, 4);
            std::memcpy(&gb, &got, 4);
            const uint64_t d = uint64_t(wb > gb ? wb - gb : gb - wb);
            lw = std::max(lw, d);
            ld += d != 0;
        }
        uint64_t cur = worst.load();
        while (lw > cur && !worst.compare_exchange_weak(cur, lw)) {}
        ndiff += ld;
    });
    if (differ) *differ = ndiff.load();
    return worst.load();
}

// Closed-form test hooks (FP32 oracle functions).
void orc_eval_bsdf(size_t n, const float* wi, const float* wo, const float* nrm, float alpha2, float metallic,
                   float spec_prob, const float* albedo, float* out) {
    for (size_t i = 0; i < n; ++i) {
        const auto b = orc::Tracer<float>::evaluate_opaque_bsdf(
            wi[3 * i], wi[3 * i + 1], wi[3 * i + 2], wo[3 * i], wo[3 * i + 1], wo[3 * i + 2], nrm[3 * i],
            nrm[3 * i + 1], nrm[3 * i + 2], alpha2, metallic, spec_prob, albedo[0], albedo[1], albedo[2]);
        out[4 * i] = b.r; out[4 * i + 1] = b.g; out[4 * i + 2] = b.b; out[4 * i + 3] = b.pdf;
    }
}

// FP64 opaque BSDF value and its derivatives by albedo (own channel) and alpha2,
// for the derivative pins.
void orc_bsdf_grad_f64(const double* wi, const double* wo, const double* nrm, double alpha2, double metallic,
                       const double* albedo, double* value, double* da, double* dl) {
    using T = orc::Tracer<double>;
    const T::Bsdf e = T::evaluate_opaque_bsdf(wi[0], wi[1], wi[2], wo[0], wo[1], wo[2], nrm[0], nrm[1], nrm[2], alpha2,
                                              metallic, 0.5, albedo[0], albedo[1], albedo[2]);
    const T::BsdfGrad g = T::evaluate_opaque_bsdf_grad(wi[0], wi[1], wi[2], wo[0], wo[1], wo[2], nrm[0], nrm[1], nrm[2],
                                                       alpha2, metallic, albedo[0], albedo[1], albedo[2]);
    value[0] = e.r; value[1] = e.g; value[2] = e.b;
    for (int c = 0; c < 3; ++c) { da[c] = g.da[c]; dl[c] = g.dl[c]; }
}

float orc_fresnel(float ci, float eta_in, float eta_out) {
    return orc::Tracer<float>::dielectric_fresnel(ci, eta_in, eta_out);
}

// n directions from one stream: kind 0 cosine hemisphere, 1 GGX half vector.
void orc_sample(uint64_t seed, int kind, size_t n, float alpha2, const float* nrm, float* out) {
    orc::Tracer<float> t;
    t.sc = nullptr;
    t.stats = nullptr;
    t.rng_.seed(seed);
    for (size_t i = 0; i < n; ++i) {
        float x, y, z;
        if (kind == 0) t.sample_cosine_hemisphere(nrm[0], nrm[1], nrm[2], x, y, z);
        else t.sample_ggx_half(alpha2, nrm[0], nrm[1], nrm[2], x, y, z);
        out[3 * i] = x; out[3 * i + 1] = y; out[3 * i + 2] = z;
    }
}

}  // extern "C"
