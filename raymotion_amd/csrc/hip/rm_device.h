// Device functions of the path tracer: RNG, 4-wide BVH traversal, surface
// lookup, BSDFs, light sampling, dielectrics and the bounce loop.
//
// Each function computes what the same-named function of the reference
// computes (cited per function) with the same IEEE float expressions in the
// same evaluation order; the library is compiled without FP contraction and
// with correctly rounded division and square root, so the kernels produce the
// same per-pixel results as a host evaluation of those expressions.
#pragma once
#include <hip/hip_runtime.h>

#include "../common/rm_atan.h"
#include "../common/rm_exp.h"
#include "../common/rm_fdiv.h"
#include "../common/rm_sincos.h"
#include "rm_layout.h"

namespace rmh {

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross3(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ V3 vmin(V3 a, V3 b) { return {fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
__device__ __forceinline__ V3 vmax(V3 a, V3 b) { return {fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
__device__ __forceinline__ V3 xyz(float4 f) { return {f.x, f.y, f.z}; }
__device__ __forceinline__ float lum3(float r, float g, float b) { return 0.2126f * r + 0.7152f * g + 0.0722f * b; }
__device__ __forceinline__ int ibits(float f) { return __float_as_int(f); }

// PCG RXS-M-XS 32 seeded by splitmix64 (pathtrace.metal:93-101).
struct Rng {
    uint32_t state;
    __device__ __forceinline__ void seed(uint64_t s) {
        uint64_t z = s + 0x9E3779B97F4A7C15ULL;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        state = uint32_t(z ^ (z >> 31));
    }
    __device__ __forceinline__ float next() {
        state = state * 747796405u + 2891336453u;
        const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
        return float(((word >> 22u) ^ word) >> 8u) * 0x1p-24f;
    }
};

// Multiplier a and increment c of n steps of Rng's state recurrence
// (s -> 747796405 s + 2891336453 mod 2^32): n steps map s to a s + c.
__device__ __forceinline__ void lcg_jump(uint32_t n, uint32_t& a, uint32_t& c) {
    uint32_t acc_a = 1u, acc_c = 0u, cur_a = 747796405u, cur_c = 2891336453u;
    while (n) {
        if (n & 1u) {
            acc_a *= cur_a;
            acc_c = acc_c * cur_a + cur_c;
        }
        cur_c = (cur_a + 1u) * cur_c;
        cur_a *= cur_a;
        n >>= 1;
    }
    a = acc_a;
    c = acc_c;
}

// Position of the set bit of m with rank r (0: the lowest); r < popcount(m).
__device__ __forceinline__ int nth_set_bit(unsigned long long m, int r) {
    int pos = 0;
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) {
        const unsigned long long lo = m & ((1ull << w) - 1ull);
        const int c = __popcll(lo);
        if (r >= c) {
            r -= c;
            m >>= w;
            pos += w;
        } else {
            m = lo;
        }
    }
    return pos;
}

__device__ __forceinline__ uint64_t pixel_seed(uint64_t seed, uint32_t x, uint32_t y) {
    return seed ^ (uint64_t(y + 1) * 0x9E3779B97F4A7C15ULL) ^ (uint64_t(x + 1) * 0xBF58476D1CE4E5B9ULL);
}
// restir_spatial's salt: the low 64 bits of the 17-digit literal at pathtrace.metal:937.
constexpr uint64_t kSpatialSalt = 0x1B54A35A6F31B458ULL;

struct Hit {
    float t, u, v;
    int tri;
};

// Surface data at a hit (HitInfo, pathtrace.metal:11-23).
struct Surface {
    float t, u, v;
    int tri;
    V3 p, n_interp, n_geom;
    float ar, ag, ab, er, eg, eb, ior, rough, metallic, alpha;
    int shader;
};

struct Tally {
    uint32_t rays, nodes, tris, overflow;
};

struct Bsdf {
    float r, g, b, pdf;
};

struct DielectricSample {
    bool valid;
    float n_x, n_y, n_z, eta_scale;
    float dlw_da2;  // d log(eta_scale) / d alpha2 (rough, unclamped weights; else 0)
};

// Derivatives of evaluate_opaque_bsdf's value: each channel by its own albedo
// channel, and every channel by alpha2 (the pdf is not differentiated).
struct BsdfGrad {
    float da_r, da_g, da_b, dl_r, dl_g, dl_b;
};

// Hooks of the bounce loop that the forward leaves empty (pathtrace_t).
// visible() answers a vertex's light-sample shadow test; the forward traces it.
struct NoAdjoint {
    template <class... T>
    __device__ __forceinline__ void emission(const T&...) {}
    template <class... T>
    __device__ __forceinline__ void nee(const T&...) {}
    template <class... T>
    __device__ __forceinline__ void scatter(const T&...) {}
    template <class... T>
    __device__ __forceinline__ void dielectric(const T&...) {}
    template <class Tr, class NS>
    __device__ __forceinline__ bool visible(Tr& t, const NS& ns) {
        return ns.shadow && !t.occluded(ns.o, ns.d, ns.dist);
    }
};

// The forward's hooks plus a record of the first 64 shadow-test outcomes of
// the path, one bit per vertex with a light sample (the adjoint's primal pass).
struct VisibilityRecorder : NoAdjoint {
    uint64_t mask = 0;
    int n = 0;
    template <class Tr, class NS>
    __device__ __forceinline__ bool visible(Tr& t, const NS& ns) {
        const bool lit = ns.shadow && !t.occluded(ns.o, ns.d, ns.dist);
        if (n < 64) mask |= uint64_t(lit) << n;
        ++n;
        return lit;
    }
};

#ifndef RM_LEAF_MAX_AGE
#define RM_LEAF_MAX_AGE 3
#endif
constexpr int kLeafMaxAge = RM_LEAF_MAX_AGE;  // traversal iterations a lane waits for a leaf batch at most
#ifndef RM_LEAF_FRAC8
#define RM_LEAF_FRAC8 4
#endif
constexpr int kLeafFrac8 = RM_LEAF_FRAC8;  // waiting share of the traversing lanes, in eighths, that runs it

// LDS copy of the light table (kLightWords float4 per light) and its CDF, for
// the kernels that stage it (Tracer::stage_lights). Accessed by name so the
// reads stay LDS reads.
__shared__ float4 lds_light_rec[kLdsLights * kLightWords];
__shared__ float lds_light_cdf[kLdsLights];

// Entries of a lane's traversal stack kept in LDS: all of them up to
// kLdsStack; the deep instance (S = kSpillStack, for trees whose stack bound
// exceeds kLdsStack) keeps its first kLdsStack entries there and the rest in
// a per-lane array in scratch memory.
template <int S>
constexpr int lds_stack_entries() { return S < kLdsStack ? S : kLdsStack; }

// Per-lane tracer state. S is the capacity of the lane's traversal stack;
// stk points at this lane's column of a [lds_stack_entries<S>()][kBlock]
// shared array.
template <int S>
struct Tracer {
    const DevScene& sc;
    const FrameParams& P;
    uint32_t* stk;
    // entries kLdsStack .. S - 1 of the deep instance (unused otherwise)
    uint32_t deep[S > kLdsStack ? S - kLdsStack : 1];
    Rng rng;
    Tally tl;

    // The light table and its CDF, copied into LDS by stage_lights when the
    // scene has at most kLdsLights lights (lds_on), else read from the scene.
    // The light sampling loops read them up to 32 times per shading point,
    // each read a dependent chain (CDF search, then the light's record).
    bool lds_on = false;

    __device__ Tracer(const DevScene& s, const FrameParams& p, uint32_t* stack) : sc(s), P(p), stk(stack) {
        rng.state = 0;
        tl = {0, 0, 0, 0};
    }
    __device__ __forceinline__ float urand() { return rng.next(); }

    // Block-wide: every thread of the block calls it once, before any other
    // use of the tracer.
    __device__ void stage_lights() {
        lds_on = sc.nlights > 0 && sc.nlights <= kLdsLights;
        if (lds_on) {
            for (int i = int(threadIdx.x); i < sc.nlights * kLightWords; i += int(blockDim.x))
                lds_light_rec[i] = sc.lights[i];
            for (int i = int(threadIdx.x); i < sc.nlights; i += int(blockDim.x)) lds_light_cdf[i] = sc.cdf[i];
        }
        __syncthreads();
    }

    struct LightRec {
        float4 a, b, c, d;  // the kLightWords words of the light's record
    };
    __device__ __forceinline__ LightRec light(int li) const {
        if (lds_on) {
            const int w = li * kLightWords;
            const LightRec r{lds_light_rec[w], lds_light_rec[w + 1], lds_light_rec[w + 2], lds_light_rec[w + 3]};
            // Keeps the optimizer from merging these reads with the global
            // branch's into one flat load through a selected pointer.
            __builtin_amdgcn_sched_barrier(0);
            return r;
        }
        const float4* L = sc.lights + size_t(li) * kLightWords;
        return LightRec{L[0], L[1], L[2], L[3]};
    }

    // ------------------------------------------------------------ traversal
    // Resumable traversal of one ray (intersect, k_persistent): the ray, the
    // closest hit so far, the node to visit next, and the sorted child slots
    // of the last node while its hit leaves wait (pend) for a leaf batch.
    struct Trav {
        V3 ro, rd;
        float ix, iy, iz, ox, oy, oz;
        Hit h;
        int node, sp;
        bool any, pend, last;
        float tn[kWide];
        int ref[kWide], kind[kWide];
    };
    // Starts a ray; false when there is nothing to traverse (T.h is the miss).
    __device__ __forceinline__ bool trav_begin(Trav& T, V3 ro, V3 rd, float tmax, bool any) {
        T.h.t = tmax;
        T.h.tri = -1;
        T.h.u = 0.0f;
        T.h.v = 0.0f;
        ++tl.rays;
        T.ro = ro;
        T.rd = rd;
        T.any = any;
        T.pend = false;
        T.last = false;
        T.sp = 0;
        T.node = 0;
        if (sc.nnodes == 0) return false;
        T.ix = 1.0f / (rd.x == 0.0f ? 1e-30f : rd.x);
        T.iy = 1.0f / (rd.y == 0.0f ? 1e-30f : rd.y);
        T.iz = 1.0f / (rd.z == 0.0f ? 1e-30f : rd.z);
        T.ox = -ro.x * T.ix;
        T.oy = -ro.y * T.iy;
        T.oz = -ro.z * T.iz;
        return true;
    }
    // Visits T.node: tests its four child boxes, sorts the hit slots by entry
    // distance, marks hit leaf slots as pending, stacks the inner children but
    // the nearest, which becomes T.node; sets T.last when nothing is left.
    __device__ __forceinline__ void trav_node(Trav& T) {
        const float4* n = sc.nodes + size_t(T.node) * kNodeWords;
        float4 w[kNodeWords];
        // Each lane loads, per axis, the near-plane word (lower bounds for a
        // positive direction component, upper bounds for a negative one) into
        // the lo slot and the far-plane word into the hi slot: with a fixed
        // sign of the inverse direction, fma(lo, i, o) <= fma(hi, i, o) or the
        // reverse holds for every non-empty child box (rounding is monotonic),
        // so the slab minima and maxima below are known without min / max.
        static_assert(kGroups == 1, "one group of four children");
        const int sx = int(__float_as_uint(T.ix) >> 31), sy = int(__float_as_uint(T.iy) >> 31),
                  sz = int(__float_as_uint(T.iz) >> 31);
        w[node_lo_word(0, 0)] = n[node_lo_word(0, 0) + sx];
        w[node_hi_word(0, 0)] = n[node_hi_word(0, 0) - sx];
        w[node_lo_word(1, 0)] = n[node_lo_word(1, 0) + sy];
        w[node_hi_word(1, 0)] = n[node_hi_word(1, 0) - sy];
        w[node_lo_word(2, 0)] = n[node_lo_word(2, 0) + sz];
        w[node_hi_word(2, 0)] = n[node_hi_word(2, 0) - sz];
        w[kNodeRefWord] = n[kNodeRefWord];
        w[kNodeKindWord] = n[kNodeKindWord];
        // The stack's top entry, read while the node is in flight: it is the
        // next node when this one has no inner child to descend into, and then
        // nothing is pushed below (entry 0 of an empty stack: unused).
        const int popped = int(stack_get(min(max(T.sp - 1, 0), S - 1)));
        ++tl.nodes;
        const float tc = T.h.t + T.h.t * (1.0f / 4096.0f);
        float* tn = T.tn;
        int* ref = T.ref;
        int* kind = T.kind;
#pragma unroll
        for (int g = 0; g < kGroups; ++g) {
            const float4 lox = w[node_lo_word(0, g)], hix = w[node_hi_word(0, g)];
            const float4 loy = w[node_lo_word(1, g)], hiy = w[node_hi_word(1, g)];
            const float4 loz = w[node_lo_word(2, g)], hiz = w[node_hi_word(2, g)];
            const float4 rf = w[kNodeRefWord + g], kd = w[kNodeKindWord + g];
            const float lx[4] = {lox.x, lox.y, lox.z, lox.w}, hx[4] = {hix.x, hix.y, hix.z, hix.w};
            const float ly[4] = {loy.x, loy.y, loy.z, loy.w}, hy[4] = {hiy.x, hiy.y, hiy.z, hiy.w};
            const float lz[4] = {loz.x, loz.y, loz.z, loz.w}, hz[4] = {hiz.x, hiz.y, hiz.z, hiz.w};
            const int r4[4] = {ibits(rf.x), ibits(rf.y), ibits(rf.z), ibits(rf.w)};
            const int k4[4] = {ibits(kd.x), ibits(kd.y), ibits(kd.z), ibits(kd.w)};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int c = 4 * g + j;
                const float ax = __builtin_fmaf(lx[j], T.ix, T.ox), bx = __builtin_fmaf(hx[j], T.ix, T.ox);
                const float ay = __builtin_fmaf(ly[j], T.iy, T.oy), by = __builtin_fmaf(hy[j], T.iy, T.oy);
                const float az = __builtin_fmaf(lz[j], T.iz, T.oz), bz = __builtin_fmaf(hz[j], T.iz, T.oz);
                // a* is the near plane's distance and b* the far plane's (an
                // empty slot, lo = +inf and hi = -inf, fails either way): the
                // values min(a, b) and max(a, b) would give
                const float nr = fmaxf(fmaxf(ax, ay), fmaxf(az, 0.0f));
                const float fr = fminf(fminf(bx, by), fminf(bz, tc));
                const bool ok = k4[j] >= 0 && nr <= fr;
                tn[c] = ok ? nr : __builtin_inff();
                ref[c] = r4[j];
                kind[c] = ok ? k4[j] : -1;
            }
        }
        // Sort the slots by entry distance (misses last): the optimal
        // 5- and 19-comparator networks for four and eight inputs.
        // Each compare-exchange as selects (no branch): the same swaps.
#define RM_CX(a, b)                                                                        \
    {                                                                                      \
        const bool sw = tn[b] < tn[a];                                                     \
        const float ta = tn[a], tb = tn[b];                                                \
        const int ra = ref[a], rb = ref[b], ka = kind[a], kb = kind[b];                    \
        tn[a] = sw ? tb : ta; tn[b] = sw ? ta : tb;                                        \
        ref[a] = sw ? rb : ra; ref[b] = sw ? ra : rb;                                      \
        kind[a] = sw ? kb : ka; kind[b] = sw ? ka : kb;                                    \
    }
        if constexpr (kWide == 4) {
            RM_CX(0, 1) RM_CX(2, 3) RM_CX(0, 2) RM_CX(1, 3) RM_CX(1, 2)
        } else {
            RM_CX(0, 2) RM_CX(1, 3) RM_CX(4, 6) RM_CX(5, 7) RM_CX(0, 4) RM_CX(1, 5) RM_CX(2, 6) RM_CX(3, 7)
            RM_CX(0, 1) RM_CX(2, 3) RM_CX(4, 5) RM_CX(6, 7) RM_CX(2, 4) RM_CX(3, 5) RM_CX(1, 4) RM_CX(3, 6)
            RM_CX(1, 2) RM_CX(3, 4) RM_CX(5, 6)
        }
#undef RM_CX
#pragma unroll
        for (int c = 0; c < kWide; ++c) T.pend = T.pend || kind[c] > 0;
        // Inner children: continue with the nearest, stack the rest (farthest
        // first). Without branches: every step stores `next` at the first free
        // entry (above the live ones, so a store that is not a push is
        // overwritten by the next one) and advances the stack only on a push.
        // The instance has more entries than the tree's stack bound (launchers
        // pick S > bound), so that entry exists; the index is clamped anyway.
        int next = -1;
#pragma unroll
        for (int c = kWide - 1; c >= 0; --c) {
            const bool inner = kind[c] == 0;
            const bool push = inner && next >= 0;
            stack_put(min(T.sp, S - 1), uint32_t(next));
            T.sp += push ? 1 : 0;
            next = inner ? ref[c] : next;
        }
        if (T.sp >= S) tl.overflow = 1;  // only a tree deeper than its stated bound
        // No inner child: the top of the stack, or the end.
        const bool pop = next < 0 && T.sp > 0;
        T.last = next < 0 && T.sp == 0;
        T.sp -= pop ? 1 : 0;
        T.node = pop ? popped : next;
    }
    __device__ __forceinline__ void stack_put(int i, uint32_t v) {
        if constexpr (S > kLdsStack) {
            if (i >= kLdsStack) {
                deep[i - kLdsStack] = v;
                return;
            }
        }
        stk[i * kBlock] = v;
    }
    __device__ __forceinline__ uint32_t stack_get(int i) const {
        if constexpr (S > kLdsStack) {
            if (i >= kLdsStack) return deep[i - kLdsStack];
        }
        return stk[i * kBlock];
    }
    // Intersects the pending leaf slots of the last node, nearest first, each
    // while its entry lies within the current hit (plus 2^-12 relative).
    // Returns true when an any-hit ray has found its hit.
    __device__ __forceinline__ bool trav_leaves(Trav& T) {
        T.pend = false;
#pragma unroll
        for (int c = 0; c < kWide; ++c) {
            if (T.kind[c] > 0 && T.tn[c] <= T.h.t + T.h.t * (1.0f / 4096.0f)) {
                const int first = T.ref[c], count = T.kind[c];
                for (int k = 0; k < count; ++k) {
                    const float4* g = sc.geo + size_t(first + k) * kTriGeoWords;
                    const float4 g0 = g[0], g1 = g[1], g2 = g[2];
                    ++tl.tris;
                    const V3 v0{g0.x, g0.y, g0.z}, v1{g0.w, g1.x, g1.y}, v2{g1.z, g1.w, g2.x};
                    const int id = ibits(g2.y);
                    const V3 e1 = v1 - v0, e2 = v2 - v0;
                    const V3 p = cross3(T.rd, e2);
                    const float det = dot3(e1, p);
                    if (fabsf(det) < 1e-12f) continue;
                    const V3 tv = T.ro - v0;
                    const float u = dot3(tv, p) / det;
                    if (u < 0.0f || u > 1.0f) continue;
                    const V3 q = cross3(tv, e1);
                    const float v = dot3(T.rd, q) / det;
                    if (v < 0.0f || u + v > 1.0f) continue;
                    const float d = dot3(e2, q) / det;
                    if (d > 1e-8f && (d < T.h.t || (d == T.h.t && T.h.tri >= 0 && id < T.h.tri))) {
                        T.h.t = d;
                        T.h.tri = id;
                        T.h.u = u;
                        T.h.v = v;
                        if (T.any) return true;
                    }
                }
            }
        }
        return false;
    }
    // Whether the wave runs a leaf batch now, for the lanes `trav` that are
    // traversing and `wait` of them whose leaves are pending: when kLeafFrac8
    // eighths of the traversing lanes wait (so also when none can go on) or
    // they have waited kLeafMaxAge iterations (age is the wave's counter).
    static __device__ __forceinline__ bool leaf_batch_now(unsigned long long wait, unsigned long long trav,
                                                          int& age, int max_age = kLeafMaxAge,
                                                          int frac8 = kLeafFrac8) {
        if (!wait) return false;
        if (8 * __popcll(wait) >= frac8 * __popcll(trav) || ++age >= max_age) {
            age = 0;
            return true;
        }
        return false;
    }

    // Closest hit (or, with any, the first hit) along o + t d for t in
    // (1e-8, tmax). Children of each 4-wide node are tested together; hit leaves
    // are intersected nearest-first and hit inner children are descended
    // nearest-first with the others kept on the LDS stack. A box is skipped
    // when its entry lies beyond the current hit by more than 2^-12 (relative),
    // and equal distances resolve to the smaller triangle index, so the result
    // is the closest hit regardless of visit order (same rule as the oracle).
    // A lane whose node has hit leaf slots waits (pend) until the wave
    // intersects leaves for all waiting lanes together (leaf_batch_now); each
    // lane's own sequence (the node's boxes, its leaves with the bound of that
    // moment, then the next node) is unchanged.
    __device__ __forceinline__ Hit intersect(V3 ro, V3 rd, float tmax, bool any) {
        Trav T;
        if (!trav_begin(T, ro, rd, tmax, any)) return T.h;
        int age = 0;
        for (;;) {
            if (!T.pend) trav_node(T);
            if (leaf_batch_now(__ballot(T.pend), __ballot(true), age) && T.pend && trav_leaves(T)) return T.h;
            if (T.last && !T.pend) break;
        }
        return T.h;
    }

    // cast_ray (pathtrace.metal:165-204): hit + interpolated normal + material.
    __device__ Surface cast_ray(V3 ro, V3 rd, float dist) { return surface_at(intersect(ro, rd, dist, false), ro, rd); }

    // Material alpha of a triangle (coverage test of cast_alpha_ray).
    __device__ __forceinline__ float tri_alpha(int tri) const {
        return sc.mats[size_t(sc.matid[tri]) * kMatWords].w;
    }

    // The surface data cast_ray derives from a hit of the ray ro + t rd.
    __device__ Surface surface_at(const Hit& h, V3 ro, V3 rd) const {
        Surface s;
        s.t = h.t; s.u = h.u; s.v = h.v; s.tri = h.tri;
        s.p = ro + rd * h.t;
        s.ar = s.ag = s.ab = s.er = s.eg = s.eb = s.ior = s.rough = s.metallic = 0.0f;
        s.alpha = 1.0f;
        s.shader = 0;
        s.n_interp = {0.0f, 0.0f, 0.0f};
        s.n_geom = {0.0f, 0.0f, 0.0f};
        if (h.tri >= 0) {
            const float4* nr = sc.nrm + size_t(h.tri) * kTriNrmWords;
            const float4 a = nr[0], b = nr[1], c = nr[2];
            const float w = 1.0f - (h.u + h.v);
            const float nx = w * a.x + (h.u * b.x + h.v * c.x);
            const float ny = w * a.y + (h.u * b.y + h.v * c.y);
            const float nz = w * a.z + (h.u * b.z + h.v * c.z);
            const float inv = 1.0f / sqrtf(nx * nx + ny * ny + nz * nz);
            s.n_interp = {nx * inv, ny * inv, nz * inv};
            s.n_geom = {a.w, b.w, c.w};
            const float4* m = sc.mats + size_t(sc.matid[h.tri]) * kMatWords;
            const float4 m0 = m[0], m1 = m[1], m2 = m[2];
            s.ar = m0.x; s.ag = m0.y; s.ab = m0.z; s.alpha = m0.w;
            s.er = m1.x; s.eg = m1.y; s.eb = m1.z; s.ior = m1.w;
            s.rough = m2.x; s.metallic = m2.y; s.shader = ibits(m2.z);
            const int texoff = ibits(m2.w);
            if (texoff >= 0) {
                const float4 m3 = m[3];
                const int tw = ibits(m3.x), th = ibits(m3.y);
                const float4 uv0 = sc.uv[size_t(h.tri) * 2], uv1 = sc.uv[size_t(h.tri) * 2 + 1];
                float fu = w * uv0.x + h.u * uv0.z + h.v * uv1.x;
                float fv = w * uv0.y + h.u * uv0.w + h.v * uv1.y;
                fu = fminf(fu - floorf(fu), 0.99999994f);
                fv = fminf(fv - floorf(fv), 0.99999994f);
                const int x = min(int(fu * float(tw)), tw - 1);
                const int y = th - 1 - min(int(fv * float(th)), th - 1);
                const float* t = sc.texels + 3 * size_t(texoff + y * tw + x);
                s.ar *= t[0]; s.ag *= t[1]; s.ab *= t[2];
            }
            if (s.shader == 1) {
                // Hue from azimuth through hsv2rgb(h, 0.8, 1) (pathtrace.metal:27-64,193-201).
                const float kpi = 3.14159265358979323846f;
                const float az = rm_atanf(s.p.x / s.p.z) * 180.0f / kpi + ((s.p.z > 0.0f) ? 0.0f : 180.0f);
                const float hh = (az - floorf(az / 360.0f) * 360.0f) / 360.0f;
                const float cc = 0.8f * 1.0f;
                const float q = hh / 60.0f;
                const float xx = cc * (1.0f - fabsf((q - floorf(q / 2.0f) * 2.0f) - 1.0f));
                const float mm = 1.0f - cc;
                float cr = 0.0f, cg = 0.0f, cb = 0.0f;
                if (hh < 60.0f) { cr = cc; cg = xx; }
                else if (hh < 120.0f) { cr = xx; cg = cc; }
                else if (hh < 180.0f) { cg = cc; cb = xx; }
                else if (hh < 240.0f) { cg = xx; cb = cc; }
                else if (hh < 300.0f) { cr = xx; cb = cc; }
                else if (hh < 360.0f) { cr = cc; cb = xx; }
                s.ar = (cr + mm) * 255.0f; s.ag = (cg + mm) * 255.0f; s.ab = (cb + mm) * 255.0f;
            }
        }
        return s;
    }

    // cast_alpha_ray (pathtrace.metal:536-550): stochastic coverage.
    __device__ Surface cast_alpha_ray(V3 ro, V3 rd, float dist) {
        if (sc.opaque) return cast_ray(ro, rd, dist);
        float traveled = 0.0f;
        while (traveled < dist) {
            Surface s = cast_ray(ro + rd * traveled, rd, dist - traveled);
            if (s.tri < 0) return s;
            const float opacity = s.alpha;
            if (opacity >= 1.0f || (opacity > 0.0f && urand() < opacity)) {
                s.t += traveled;
                return s;
            }
            traveled += s.t + 0.0001f;
        }
        Surface miss;
        miss.t = 1e30f; miss.tri = -1; miss.u = miss.v = 0.0f;
        miss.ar = miss.ag = miss.ab = miss.er = miss.eg = miss.eb = miss.ior = miss.rough = miss.metallic = 0.0f;
        miss.alpha = 1.0f; miss.shader = 0;
        miss.p = miss.n_interp = miss.n_geom = {0.0f, 0.0f, 0.0f};
        return miss;
    }

    __device__ bool occluded(V3 o, V3 d, float dist) {
        if (sc.opaque) return intersect(o, d, dist, true).tri >= 0;
        return cast_alpha_ray(o, d, dist).tri >= 0;
    }

    // ------------------------------------------------------------ sampling
    static __device__ __forceinline__ void make_tangent(float nx, float ny, float nz, float& tx, float& ty, float& tz,
                                                        float& bx, float& by, float& bz) {
        if (fabsf(nx) > 0.1f) { tx = nz; ty = 0.0f; tz = -nx; }
        else { tx = 0.0f; ty = -nz; tz = ny; }
        const float l = 1.0f / sqrtf(tx * tx + ty * ty + tz * tz);
        tx *= l; ty *= l; tz *= l;
        bx = ny * tz - nz * ty;
        by = nz * tx - nx * tz;
        bz = nx * ty - ny * tx;
    }

    // sample_cosine_hemisphere (pathtrace.metal:205-220)
    __device__ void sample_cosine(float nx, float ny, float nz, float& ox, float& oy, float& oz) {
        const float ru = urand();
        const float phi = urand() * 6.2831855f;
        float sp, c;
        rm_sincosf(phi, &sp, &c);
        const float rr = sqrtf(ru);
        const float lx = rr * c, ly = rr * sp, lz = sqrtf(fmaxf(0.0f, 1.0f - ru));
        float tx, ty, tz, bx, by, bz;
        make_tangent(nx, ny, nz, tx, ty, tz, bx, by, bz);
        ox = tx * lx + (bx * ly + nx * lz);
        oy = ty * lx + (by * ly + ny * lz);
        oz = tz * lx + (bz * ly + nz * lz);
    }

    // sample_ggx_half (pathtrace.metal:221-242); its first draw is unused.
    __device__ void sample_ggx_half(float alpha2, float nx, float ny, float nz, float& hx, float& hy, float& hz) {
        (void)urand();
        const float rv = urand();
        const float phi = urand() * 6.2831855f;
        float sp, c;
        rm_sincosf(phi, &sp, &c);
        const float cs = sqrtf((1.0f - rv) / (1.0f + (alpha2 - 1.0f) * rv));
        const float sn = sqrtf(fmaxf(0.0f, 1.0f - cs * cs));
        const float lx = sn * c, ly = sn * sp, lz = cs;
        float tx, ty, tz, bx, by, bz;
        make_tangent(nx, ny, nz, tx, ty, tz, bx, by, bz);
        hx = tx * lx + (bx * ly + nx * lz);
        hy = ty * lx + (by * ly + ny * lz);
        hz = tz * lx + (bz * ly + nz * lz);
        const float d = 1.0f / sqrtf(hx * hx + hy * hy + hz * hz);
        hx *= d; hy *= d; hz *= d;
    }

    // evaluate_opaque_bsdf (pathtrace.metal:243-295): Lambert + GGX/Smith/Schlick.
    // Derivatives of eval_bsdf's value by albedo and alpha2, from the same
    // intermediate quantities: Schlick F' = m (1 - om); GGX
    // D' = (dd - 2 alpha2 cos_h^2) / (pi dd^3); Smith G1' = -c (1 - c^2) / (s (c + s)^2)
    // with s = sqrt(c^2 + alpha2 (1 - c^2)).
    static __device__ __forceinline__ BsdfGrad eval_bsdf_grad(float wi_x, float wi_y, float wi_z, float wo_x,
                                                              float wo_y, float wo_z, float n_x, float n_y, float n_z,
                                                              float alpha2, float metallic, float ar, float ag,
                                                              float ab) {
        const float kpi = 3.14159265358979323846f;
        BsdfGrad out{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        const float cos_i = wi_x * n_x + (wi_y * n_y + wi_z * n_z);
        const float cos_o = wo_x * n_x + (wo_y * n_y + wo_z * n_z);
        if (!(cos_i > 1e-8f && cos_o > 1e-8f)) return out;
        float hx = wi_x + wo_x, hy = wi_y + wo_y, hz = wi_z + wo_z;
        const float hl = sqrtf(hx * hx + hy * hy + hz * hz);
        if (!(hl > 1e-8f)) return out;
        const float inv = 1.0f / hl;
        hx *= inv; hy *= inv; hz *= inv;
        float cos_h = hx * n_x + (hy * n_y + hz * n_z);
        float vo_h = wo_x * hx + (wo_y * hy + wo_z * hz);
        if (cos_h < 0.0f) { cos_h = -cos_h; vo_h = -vo_h; }
        if (!(vo_h > 1e-8f)) return out;
        const float dd = cos_h * cos_h * (alpha2 - 1.0f) + 1.0f;
        const float D = alpha2 / (kpi * dd * dd);
        const float so = sqrtf(cos_o * cos_o + alpha2 * (1.0f - cos_o * cos_o));
        const float si = sqrtf(cos_i * cos_i + alpha2 * (1.0f - cos_i * cos_i));
        const float g1 = 2.0f * cos_o / (cos_o + so);
        const float g2 = 2.0f * cos_i / (cos_i + si);
        const float G = g1 * g2;
        float om = 1.0f - vo_h;
        om *= om;
        om *= om;
        om *= 1.0f - vo_h;
        const float di = 1.0f - metallic;
        const float f0r = 0.04f * di + ar * metallic, f0g = 0.04f * di + ag * metallic, f0b = 0.04f * di + ab * metallic;
        const float fr = f0r + (1.0f - f0r) * om, fg = f0g + (1.0f - f0g) * om, fb = f0b + (1.0f - f0b) * om;
        const float base = D * G / (4.0f * cos_i * cos_o);
        const float dF = metallic * (1.0f - om);
        out.da_r = rm_div_pi(di * (1.0f - fr) - ar * di * dF) + base * dF;
        out.da_g = rm_div_pi(di * (1.0f - fg) - ag * di * dF) + base * dF;
        out.da_b = rm_div_pi(di * (1.0f - fb) - ab * di * dF) + base * dF;
        const float dD = (dd - 2.0f * alpha2 * cos_h * cos_h) / (kpi * dd * dd * dd);
        const float dg1 = -cos_o * (1.0f - cos_o * cos_o) / (so * (cos_o + so) * (cos_o + so));
        const float dg2 = -cos_i * (1.0f - cos_i * cos_i) / (si * (cos_i + si) * (cos_i + si));
        const float dbase = (dD * G + D * (dg1 * g2 + g1 * dg2)) / (4.0f * cos_i * cos_o);
        out.dl_r = dbase * fr;
        out.dl_g = dbase * fg;
        out.dl_b = dbase * fb;
        return out;
    }

    static __device__ __forceinline__ Bsdf eval_bsdf(float wi_x, float wi_y, float wi_z, float wo_x, float wo_y, float wo_z, float n_x,
                                     float n_y, float n_z, float alpha2, float metallic, float spec_prob, float ar,
                                     float ag, float ab) {
        const float kpi = 3.14159265358979323846f;
        Bsdf out{0.0f, 0.0f, 0.0f, 0.0f};
        const float cos_i = wi_x * n_x + (wi_y * n_y + wi_z * n_z);
        const float cos_o = wo_x * n_x + (wo_y * n_y + wo_z * n_z);
        if (!(cos_i > 1e-8f && cos_o > 1e-8f)) return out;
        float hx = wi_x + wo_x, hy = wi_y + wo_y, hz = wi_z + wo_z;
        const float hl = sqrtf(hx * hx + hy * hy + hz * hz);
        if (!(hl > 1e-8f)) return out;
        const float inv = 1.0f / hl;
        hx *= inv; hy *= inv; hz *= inv;
        float cos_h = hx * n_x + (hy * n_y + hz * n_z);
        float vo_h = wo_x * hx + (wo_y * hy + wo_z * hz);
        if (cos_h < 0.0f) { cos_h = -cos_h; vo_h = -vo_h; }
        if (!(vo_h > 1e-8f)) return out;
        const float dd = cos_h * cos_h * (alpha2 - 1.0f) + 1.0f;
        const float D = alpha2 / (kpi * dd * dd);
        const float g1 = 2.0f * cos_o / (cos_o + sqrtf(cos_o * cos_o + alpha2 * (1.0f - cos_o * cos_o)));
        const float g2 = 2.0f * cos_i / (cos_i + sqrtf(cos_i * cos_i + alpha2 * (1.0f - cos_i * cos_i)));
        const float G = g1 * g2;
        float om = 1.0f - vo_h;
        om *= om;
        om *= om;
        om *= 1.0f - vo_h;
        const float di = 1.0f - metallic;
        const float f0r = 0.04f * di + ar * metallic, f0g = 0.04f * di + ag * metallic, f0b = 0.04f * di + ab * metallic;
        const float fr = f0r + (1.0f - f0r) * om, fg = f0g + (1.0f - f0g) * om, fb = f0b + (1.0f - f0b) * om;
        const float kdr = ar * di * (1.0f - fr), kdg = ag * di * (1.0f - fg), kdb = ab * di * (1.0f - fb);
        const float base = D * G / (4.0f * cos_i * cos_o);
        out.r = rm_div_pi(kdr) + base * fr;
        out.g = rm_div_pi(kdg) + base * fg;
        out.b = rm_div_pi(kdb) + base * fb;
        const float pdf_d = rm_div_pi(cos_i);
        const float pdf_s = D * cos_h / (4.0f * vo_h);
        out.pdf = (1.0f - spec_prob) * pdf_d + spec_prob * pdf_s;
        return out;
    }

    // CDF binary search (pathtrace.metal:323-330), over the LDS copy when staged.
    __device__ int pick_light(float pick) const {
        int lo = 0, hi = sc.nlights - 1;
        if (lds_on) {
            while (lo < hi) {
                const int mid = (lo + hi) / 2;
                if (lds_light_cdf[mid] >= pick) hi = mid;
                else lo = mid + 1;
            }
        } else {
            while (lo < hi) {
                const int mid = (lo + hi) / 2;
                if (sc.cdf[mid] >= pick) hi = mid;
                else lo = mid + 1;
            }
        }
        return lo;
    }

    // direct_lighting (pathtrace.metal:299-469) up to its shadow ray: legacy
    // NEE for one candidate, otherwise RIS over Mc candidates. Consumes every
    // RNG draw the reference takes before the shadow ray; the caller traces
    // the ray (if any) and keeps (r, g, b) only when it is unoccluded.
    struct NeeSample {
        bool shadow;
        V3 o, d;
        float dist;
        float r, g, b, sumW;
        uint32_t M;
        // Legacy NEE with a contribution: light index (CDF order), its
        // emission, the BSDF value toward it and the factor k with
        // (r, g, b) = emission * bsdf * k. Read only by the adjoint.
        int light;
        float ler, leg, leb, er, eg, eb, k;
    };
    // direct_lighting's RIS (pathtrace.metal:299-469): the shading point,
    // one candidate, and the running reservoir of candidates taken so far.
    struct RisIn {
        float p_x, p_y, p_z, n_x, n_y, n_z, wo_x, wo_y, wo_z, alpha2, metallic, spec_prob, ar, ag, ab;
    };
    struct RisCand {
        float qx, qy, qz, w, ph, epdf, Fr, Fg, Fb;
    };
    struct RisPick {
        float sumW = 0.0f;
        float lx = 0.0f, ly = 0.0f, lz = 0.0f, Fr = 0.0f, Fg = 0.0f, Fb = 0.0f, phat = 0.0f, epdf = 0.0f;
        bool have = false;
    };
    // Candidate at folded barycentrics (lu, lv) of light record lr.
    __device__ __forceinline__ RisCand ris_candidate(const LightRec& lr, float lu, float lv, const RisIn& in) const {
        const float4 l0 = lr.a, l1 = lr.b, l2 = lr.c, l3 = lr.d;
        if (lu + lv > 1.0f) { lu = 1.0f - lu; lv = 1.0f - lv; }
        const float lw = 1.0f - lu - lv;
        RisCand c;
        c.qx = lw * l0.x + (lu * l0.w + lv * l1.z);
        c.qy = lw * l0.y + (lu * l1.x + lv * l1.w);
        c.qz = lw * l0.z + (lu * l1.y + lv * l2.x);
        c.w = c.ph = c.epdf = c.Fr = c.Fg = c.Fb = 0.0f;
        const float ldx = c.qx - in.p_x, ldy = c.qy - in.p_y, ldz = c.qz - in.p_z;
        const float dist = sqrtf(ldx * ldx + ldy * ldy + ldz * ldz);
        if (dist > 0.003f) {
            const float inv = 1.0f / dist;
            const float dxn = ldx * inv, dyn = ldy * inv, dzn = ldz * inv;
            const float lcos = -(l2.y * dxn + (l2.z * dyn + l2.w * dzn));
            const float scos = in.n_x * dxn + (in.n_y * dyn + in.n_z * dzn);
            if (lcos > 1e-8f && scos > 1e-8f) {
                float lt_lum = lum3(l3.x, l3.y, l3.z);
                if (lt_lum < 1e-6f) lt_lum = 1e-6f;
                const float q = dist * dist * lt_lum / (lcos * sc.light_total);
                if (q > 1e-12f) {
                    const Bsdf e = eval_bsdf(dxn, dyn, dzn, in.wo_x, in.wo_y, in.wo_z, in.n_x, in.n_y, in.n_z,
                                             in.alpha2, in.metallic, in.spec_prob, in.ar, in.ag, in.ab);
                    c.epdf = e.pdf;
                    c.Fr = l3.x * e.r * scos * l3.w;
                    c.Fg = l3.y * e.g * scos * l3.w;
                    c.Fb = l3.z * e.b * scos * l3.w;
                    c.ph = lum3(c.Fr, c.Fg, c.Fb);
                    if (c.ph > 0.0f) c.w = c.ph / q;
                }
            }
        }
        return c;
    }
    // Reservoir update with the candidate's draw r.
    static __device__ __forceinline__ void ris_take(RisPick& s, const RisCand& c, float r) {
        s.sumW += c.w;
        if (c.w > 0.0f && r * s.sumW < c.w) {
            s.have = true;
            s.lx = c.qx; s.ly = c.qy; s.lz = c.qz;
            s.Fr = c.Fr; s.Fg = c.Fg; s.Fb = c.Fb;
            s.phat = c.ph; s.epdf = c.epdf;
        }
    }
    // The light sample of a finished reservoir: its shadow ray and weight.
    __device__ void ris_finish(NeeSample& ns, const RisPick& s, int Mc, const RisIn& in) const {
        ns.sumW = s.sumW;
        ns.M = uint32_t(Mc);
        if (!s.have || !(s.sumW > 0.0f) || !(s.phat > 0.0f)) return;
        const float pdf_ris = float(Mc) * s.phat / s.sumW;
        float wnee = 1.0f;
        if (P.nomis == 0 && s.epdf > 1e-20f) wnee = pdf_ris / (pdf_ris + s.epdf);
        float vdx = s.lx - in.p_x, vdy = s.ly - in.p_y, vdz = s.lz - in.p_z;
        const float vdist = sqrtf(vdx * vdx + vdy * vdy + vdz * vdz);
        if (!(vdist > 0.003f)) return;
        const float vinv = 1.0f / vdist;
        vdx *= vinv; vdy *= vinv; vdz *= vinv;
        const float W = s.sumW / (float(Mc) * s.phat);
        ns.shadow = true;
        ns.o = V3{in.p_x + in.n_x * 0.002f, in.p_y + in.n_y * 0.002f, in.p_z + in.n_z * 0.002f};
        ns.d = V3{vdx, vdy, vdz};
        ns.dist = vdist - 0.003f;
        ns.r = s.Fr * W * wnee;
        ns.g = s.Fg * W * wnee;
        ns.b = s.Fb * W * wnee;
    }

    __device__ __forceinline__ NeeSample nee_prepare(float p_x, float p_y, float p_z, float n_x, float n_y, float n_z, float wo_x,
                                     float wo_y, float wo_z, float rough, float metallic, float spec_prob, float ar,
                                     float ag, float ab) {
        NeeSample ns;
        ns.shadow = false;
        ns.o = ns.d = V3{0.0f, 0.0f, 0.0f};
        ns.dist = 0.0f;
        ns.r = ns.g = ns.b = ns.sumW = 0.0f;
        ns.M = 0;
        ns.light = -1;
        ns.ler = ns.leg = ns.leb = ns.er = ns.eg = ns.eb = ns.k = 0.0f;
        if (sc.nlights == 0 || !(sc.light_total > 1e-12f)) return ns;
        int Mc = P.restir_candidates;
        if (Mc < 1) Mc = 1;
        if (Mc > 32) Mc = 32;
        float r2 = rough;
        if (r2 < 0.0f) r2 = 0.0f;
        if (r2 > 1.0f) r2 = 1.0f;
        if (r2 < 0.02f) r2 = 0.02f;
        const float alpha2 = (r2 * r2) * (r2 * r2);
        const bool nomis = P.nomis != 0;
        const float total = sc.light_total;
        const V3 so{p_x + n_x * 0.002f, p_y + n_y * 0.002f, p_z + n_z * 0.002f};
        if (Mc <= 1) {
            const float pick = urand() * total;
            const int li = pick_light(pick);
            const LightRec lr = light(li);
            const float4 l0 = lr.a, l1 = lr.b, l2 = lr.c, l3 = lr.d;
            float lu = urand(), lv = urand();
            if (lu + lv > 1.0f) { lu = 1.0f - lu; lv = 1.0f - lv; }
            const float lw = 1.0f - lu - lv;
            float ldx = lw * l0.x + (lu * l0.w + lv * l1.z) - p_x;
            float ldy = lw * l0.y + (lu * l1.x + lv * l1.w) - p_y;
            float ldz = lw * l0.z + (lu * l1.y + lv * l2.x) - p_z;
            const float dist = sqrtf(ldx * ldx + ldy * ldy + ldz * ldz);
            ns.M = 1;
            if (!(dist > 0.003f)) return ns;
            const float inv = 1.0f / dist;
            ldx *= inv; ldy *= inv; ldz *= inv;
            const float lcos = -(l2.y * ldx + (l2.z * ldy + l2.w * ldz));
            const float scos = n_x * ldx + (n_y * ldy + n_z * ldz);
            if (!(lcos > 1e-8f && scos > 1e-8f)) return ns;
            float lt_lum = lum3(l3.x, l3.y, l3.z);
            if (lt_lum < 1e-6f) lt_lum = 1e-6f;
            const float pdf = dist * dist * lt_lum / (lcos * total);
            const Bsdf e = eval_bsdf(ldx, ldy, ldz, wo_x, wo_y, wo_z, n_x, n_y, n_z, alpha2, metallic, spec_prob, ar,
                                     ag, ab);
            const float Fr = l3.x * e.r * scos * l3.w, Fg = l3.y * e.g * scos * l3.w, Fb = l3.z * e.b * scos * l3.w;
            const float ph = lum3(Fr, Fg, Fb);
            // The same sum whether or not the shadow ray is blocked (metal:364,370,375).
            ns.sumW = (ph > 0.0f && pdf > 1e-12f) ? ph / pdf : 0.0f;
            ns.shadow = true;
            ns.o = so;
            ns.d = V3{ldx, ldy, ldz};
            ns.dist = dist - 0.003f;
            if (pdf > 1e-12f) {
                float wnee = 1.0f;
                if (!nomis && e.pdf > 1e-20f) wnee = pdf / (pdf + e.pdf);
                const float k = scos * wnee * l3.w / pdf;
                ns.r = l3.x * e.r * k;
                ns.g = l3.y * e.g * k;
                ns.b = l3.z * e.b * k;
                ns.light = li;
                ns.ler = l3.x; ns.leg = l3.y; ns.leb = l3.z;
                ns.er = e.r; ns.eg = e.g; ns.eb = e.b;
                ns.k = k;
            }
            return ns;
        }
        const RisIn in{p_x, p_y, p_z, n_x, n_y, n_z, wo_x, wo_y, wo_z, alpha2, metallic, spec_prob, ar, ag, ab};
        RisPick pk;
        // Candidates in pairs, so the two light searches and record reads
        // overlap; each candidate's draws (pick, lu, lv, then the reservoir's
        // r) and the reservoir updates keep their order.
        int j = 0;
        for (; j + 1 < Mc; j += 2) {
            const float pa = urand() * total;
            const float ua = urand(), va = urand(), ra = urand();
            const float pb = urand() * total;
            const float ub = urand(), vb = urand(), rb = urand();
            const LightRec A = light(pick_light(pa));
            const LightRec B = light(pick_light(pb));
            const RisCand ca = ris_candidate(A, ua, va, in);
            const RisCand cb = ris_candidate(B, ub, vb, in);
            ris_take(pk, ca, ra);
            ris_take(pk, cb, rb);
        }
        if (j < Mc) {
            const float pa = urand() * total;
            const float ua = urand(), va = urand(), ra = urand();
            ris_take(pk, ris_candidate(light(pick_light(pa)), ua, va, in), ra);
        }
        ris_finish(ns, pk, Mc, in);
        return ns;
    }

    // direct_lighting's RIS for all lanes of the wave in `W` at once (each
    // with its own shading point, in the arguments): every lane of the wave
    // evaluates candidates of one waiting lane (group lane % |W|). A candidate
    // takes exactly four draws (pick, lu, lv, r), so candidate j of a waiting
    // lane starts 4j steps into its random sequence, reached by jumping its
    // LCG state; each waiting lane then folds its candidates' weights in
    // candidate order, so its sums, selection and final random state are those
    // of nee_prepare bit for bit. The cost is ceil(Mc |W| / 64) candidate
    // evaluations per lane instead of Mc. Called by every lane of the wave,
    // with the same W, and only when nee_prepare would take its RIS branch
    // (ris_coop_applies).
    __device__ bool ris_coop_applies() const {
        return sc.nlights != 0 && sc.light_total > 1e-12f && P.restir_candidates > 1;
    }
    __device__ NeeSample nee_prepare_coop(unsigned long long W, float p_x, float p_y, float p_z, float n_x, float n_y,
                                          float n_z, float wo_x, float wo_y, float wo_z, float rough, float metallic,
                                          float spec_prob, float ar, float ag, float ab) {
        NeeSample ns;
        ns.shadow = false;
        ns.o = ns.d = V3{0.0f, 0.0f, 0.0f};
        ns.dist = 0.0f;
        ns.r = ns.g = ns.b = ns.sumW = 0.0f;
        ns.M = 0;
        ns.light = -1;
        ns.ler = ns.leg = ns.leb = ns.er = ns.eg = ns.eb = ns.k = 0.0f;
        int Mc = P.restir_candidates;
        if (Mc > 32) Mc = 32;
        float r2 = rough;
        if (r2 < 0.0f) r2 = 0.0f;
        if (r2 > 1.0f) r2 = 1.0f;
        if (r2 < 0.02f) r2 = 0.02f;
        const float alpha2 = (r2 * r2) * (r2 * r2);
        const float total = sc.light_total;
        const int lane = int(__lane_id());
        const bool waiting = (W >> lane) & 1ull;
        const int nw = __popcll(W);
        const int grp = lane % nw, sub = lane / nw;
        const int gsize = (64 - grp + nw - 1) / nw;       // lanes of group grp
        const int lmin = 64 / nw, lmax = (64 + nw - 1) / nw;
        const int owner = nth_set_bit(W, grp);           // the waiting lane this lane works for
        const int mine = int(__builtin_amdgcn_mbcnt_hi(unsigned(W >> 32), __builtin_amdgcn_mbcnt_lo(unsigned(W), 0u)));
        const int msize = (64 - mine + nw - 1) / nw;      // size of this lane's own group, when waiting
        const RisIn in{__shfl(p_x, owner), __shfl(p_y, owner), __shfl(p_z, owner), __shfl(n_x, owner),
                       __shfl(n_y, owner), __shfl(n_z, owner), __shfl(wo_x, owner), __shfl(wo_y, owner),
                       __shfl(wo_z, owner), __shfl(alpha2, owner), __shfl(metallic, owner),
                       __shfl(spec_prob, owner), __shfl(ar, owner), __shfl(ag, owner), __shfl(ab, owner)};
        const uint32_t s0 = uint32_t(__shfl(int(rng.state), owner));
        uint32_t ja, jc, sa, sc_;
        lcg_jump(4u * uint32_t(sub), ja, jc);     // to this lane's first candidate
        lcg_jump(4u * uint32_t(gsize), sa, sc_);  // from one of its candidates to the next
        uint32_t st = ja * s0 + jc;
        RisPick pk;
        int sel_lane = -1;
        RisCand keep{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        const int rounds = (Mc + lmin - 1) / lmin;
        const int steps = lmax < Mc ? lmax : Mc;
        for (int k = 0; k < rounds; ++k) {
            const int jc0 = k * gsize + sub;  // this lane's candidate in round k
            RisCand c{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            float r = 0.0f;
            if (jc0 < Mc) {
                Rng g;
                g.state = st;
                const float pick = g.next() * total;
                const float lu = g.next(), lv = g.next();
                r = g.next();
                c = ris_candidate(light(pick_light(pick)), lu, lv, in);
            }
            st = sa * st + sc_;
            // Waiting lanes fold their group's candidates of this round in order.
            int round_sel = -1;
            for (int t = 0; t < steps; ++t) {
                const bool mine_t = waiting && t < msize && k * msize + t < Mc;
                const int src = mine_t ? mine + t * nw : lane;
                const float wt = __shfl(c.w, src), rt = __shfl(r, src);
                if (mine_t) {
                    pk.sumW += wt;
                    if (wt > 0.0f && rt * pk.sumW < wt) {
                        pk.have = true;
                        round_sel = src;
                    }
                }
            }
            // A lane keeps its candidate while its owner's selection is it.
            if (__shfl(round_sel, owner) == lane) keep = c;
            if (round_sel >= 0) sel_lane = round_sel;
        }
        const int from = sel_lane >= 0 ? sel_lane : lane;
        pk.lx = __shfl(keep.qx, from); pk.ly = __shfl(keep.qy, from); pk.lz = __shfl(keep.qz, from);
        pk.Fr = __shfl(keep.Fr, from); pk.Fg = __shfl(keep.Fg, from); pk.Fb = __shfl(keep.Fb, from);
        pk.phat = __shfl(keep.ph, from); pk.epdf = __shfl(keep.epdf, from);
        if (waiting) {
            uint32_t ma, mcn;
            lcg_jump(4u * uint32_t(Mc), ma, mcn);
            rng.state = ma * rng.state + mcn;
            const RisIn own{p_x, p_y, p_z, n_x, n_y, n_z, wo_x, wo_y, wo_z, alpha2, metallic, spec_prob, ar, ag, ab};
            ris_finish(ns, pk, Mc, own);
        }
        return ns;
    }

    static __device__ float fresnel(float ci, float eta_in, float eta_out) {  // pathtrace.metal:80-91
        if (ci < 0.0f) ci = 0.0f;
        if (ci > 1.0f) ci = 1.0f;
        const float eta = eta_in / eta_out;
        const float s2 = eta * eta * (1.0f - ci * ci);
        if (s2 >= 1.0f) return 1.0f;
        const float ct = sqrtf(1.0f - s2);
        const float rp = (eta_out * ci - eta_in * ct) / (eta_out * ci + eta_in * ct);
        const float rs = (eta_in * ci - eta_out * ct) / (eta_in * ci + eta_out * ct);
        return 0.5f * (rp * rp + rs * rs);
    }

    // sample_dielectric (pathtrace.metal:470-534)
    __device__ DielectricSample sample_dielectric(float wo_x, float wo_y, float wo_z, float n_x, float n_y, float n_z,
                                                  float cdx, float cdy, float cdz, float ior, bool front, float rough,
                                                  float alpha2) {
        DielectricSample out{true, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f};
        float eta_i = 1.0f, eta_t = ior;
        if (!front) { eta_i = ior; eta_t = 1.0f; }
        float hx, hy, hz;
        if (rough <= 0.0201f) {
            hx = n_x; hy = n_y; hz = n_z;
        } else {
            sample_ggx_half(alpha2, n_x, n_y, n_z, hx, hy, hz);
            const float md = wo_x * hx + (wo_y * hy + wo_z * hz);
            if (!(md > 1e-8f)) out.valid = false;
        }
        float cos_m = wo_x * hx + (wo_y * hy + wo_z * hz);
        if (cos_m < 0.0f) { hx = -hx; hy = -hy; hz = -hz; cos_m = -cos_m; }
        if (cos_m > 1.0f) cos_m = 1.0f;
        const float fres = fresnel(cos_m, eta_i, eta_t);
        const float eta = eta_i / eta_t;
        const float s2 = eta * eta * (1.0f - cos_m * cos_m);
        if (s2 >= 1.0f || urand() < fres) {
            out.n_x = 2.0f * cos_m * hx - wo_x;
            out.n_y = 2.0f * cos_m * hy - wo_y;
            out.n_z = 2.0f * cos_m * hz - wo_z;
            out.eta_scale = 1.0f;
        } else {
            const float cos_t = sqrtf(1.0f - s2);
            out.n_x = eta * cdx + hx * (eta * cos_m - cos_t);
            out.n_y = eta * cdy + hy * (eta * cos_m - cos_t);
            out.n_z = eta * cdz + hz * (eta * cos_m - cos_t);
            out.eta_scale = eta * eta;
        }
        {
            const float d = 1.0f / sqrtf(out.n_x * out.n_x + out.n_y * out.n_y + out.n_z * out.n_z);
            out.n_x *= d; out.n_y *= d; out.n_z *= d;
        }
        if (rough > 0.0201f) {
            const float cos_i = fabsf(out.n_x * n_x + (out.n_y * n_y + out.n_z * n_z));
            const float cos_o = fabsf(wo_x * n_x + (wo_y * n_y + wo_z * n_z));
            const float cos_h = fabsf(hx * n_x + (hy * n_y + hz * n_z));
            const float vo_h = fabsf(wo_x * hx + (wo_y * hy + wo_z * hz));
            if (cos_i > 1e-8f && cos_o > 1e-8f && cos_h > 1e-8f && vo_h > 1e-8f) {
                const float so = sqrtf(cos_o * cos_o + alpha2 * (1.0f - cos_o * cos_o));
                const float si = sqrtf(cos_i * cos_i + alpha2 * (1.0f - cos_i * cos_i));
                const float g1 = 2.0f * cos_o / (cos_o + so);
                const float g2 = 2.0f * cos_i / (cos_i + si);
                float w = g1 * g2 * vo_h / (cos_o * cos_h);
                // (log G1)' = -(1 - c^2) / (2 s (c + s)); zero where the weight is clamped
                if (!(w > 10.0f))
                    out.dlw_da2 = -(1.0f - cos_o * cos_o) / (2.0f * so * (cos_o + so)) -
                                  (1.0f - cos_i * cos_i) / (2.0f * si * (cos_i + si));
                if (w > 10.0f) w = 10.0f;
                out.eta_scale *= w;
            } else {
                out.valid = false;
            }
        }
        return out;
    }

    // MIS weight of emission reached by a BSDF-sampled ray, with the
    // ReSTIR-aware light pdf (pathtrace.metal:595-636 and 1191-1209).
    __device__ float emission_mis(const Surface& hit, float ox, float oy, float oz, float last_pdf, float prev_sumW,
                                  uint32_t prev_M, float pb_r, float pb_g, float pb_b, float pb_cos) {
        float lx = hit.p.x - ox, ly = hit.p.y - oy, lz = hit.p.z - oz;
        const float dist2 = lx * lx + ly * ly + lz * lz;
        const float dist = sqrtf(dist2);
        float pl = 1e30f;
        const float4 nn = sc.nn[hit.tri];
        if (dist > 1e-20f && last_pdf > 1e-20f && prev_M > 0 && prev_sumW > 1e-20f) {
            const float ph = lum3(hit.er * pb_r * pb_cos * hit.alpha, hit.eg * pb_g * pb_cos * hit.alpha,
                                  hit.eb * pb_b * pb_cos * hit.alpha);
            if (ph > 1e-20f) {
                pl = float(prev_M) * ph / prev_sumW;
            } else {
                const float inv = 1.0f / dist;
                const float lxn = lx * inv, lyn = ly * inv, lzn = lz * inv;
                const float cos_l = -(nn.x * lxn + (nn.y * lyn + nn.z * lzn));
                if (cos_l > 1e-8f) {
                    float lum = lum3(hit.er, hit.eg, hit.eb);
                    if (lum < 1e-6f) lum = 1e-6f;
                    pl = dist2 * lum / (cos_l * sc.light_total);
                }
            }
        } else if (dist > 1e-20f && last_pdf > 1e-20f) {
            const float inv = 1.0f / dist;
            lx *= inv; ly *= inv; lz *= inv;
            const float cos_l = -(nn.x * lx + (nn.y * ly + nn.z * lz));
            if (cos_l > 1e-8f) {
                float lum = lum3(hit.er, hit.eg, hit.eb);
                if (lum < 1e-6f) lum = 1e-6f;
                pl = dist2 * lum / (cos_l * sc.light_total);
            }
        }
        return P.nomis ? 0.0f : ((last_pdf + pl > 1e-20f) ? last_pdf / (last_pdf + pl) : 0.0f);
    }

    // pathtrace (pathtrace.metal:552-728). The hooks of `adj` see every term
    // of the estimate as it is added (the adjoint replays paths through them).
    __device__ V3 pathtrace(V3 ro, V3 rd, int max_bounces) {
        NoAdjoint none;
        return pathtrace_t(ro, rd, max_bounces, none);
    }
    template <class A>
    __device__ __forceinline__ V3 pathtrace_t(V3 ro, V3 rd, int max_bounces, A& adj) {
        float ox = ro.x, oy = ro.y, oz = ro.z, dx = rd.x, dy = rd.y, dz = rd.z;
        float accr = 0.0f, accg = 0.0f, accb = 0.0f, thr_r = 1.0f, thr_g = 1.0f, thr_b = 1.0f;
        int bounce = 0, prev_delta = 1;
        float last_pdf = 0.0f, prev_sumW = 0.0f;
        uint32_t prev_M = 0;
        float prev_br = 0.0f, prev_bg = 0.0f, prev_bb = 0.0f, prev_cos = 0.0f;
        for (int b = 0; b < max_bounces; ++b) {
            const Surface hit = cast_alpha_ray({ox, oy, oz}, {dx, dy, dz}, 1e30f);
            if (hit.tri < 0) break;
            float pnx = hit.n_interp.x, pny = hit.n_interp.y, pnz = hit.n_interp.z;
            const float geom = hit.n_geom.x * pnx + (hit.n_geom.y * pny + hit.n_geom.z * pnz);
            if (geom < 0.0f) { pnx = -pnx; pny = -pny; pnz = -pnz; }
            const float ndot = dx * pnx + (dy * pny + dz * pnz);
            bool front = true;
            if (ndot > 0.0f) { pnx = -pnx; pny = -pny; pnz = -pnz; front = false; }
            const float wox = -dx, woy = -dy, woz = -dz;
            float rough = hit.rough;
            if (rough < 0.0f) rough = 0.0f;
            if (rough > 1.0f) rough = 1.0f;
            if (rough < 0.02f) rough = 0.02f;
            const float alpha = rough * rough;
            const float alpha2 = alpha * alpha;
            float metallic = hit.metallic;
            if (metallic < 0.0f) metallic = 0.0f;
            if (metallic > 1.0f) metallic = 1.0f;
            if (bounce == 0 || prev_delta == 1) {
                accr += hit.er * thr_r; accg += hit.eg * thr_g; accb += hit.eb * thr_b;
                adj.emission(hit, thr_r, thr_g, thr_b, 1.0f);
            } else {
                const float w = emission_mis(hit, ox, oy, oz, last_pdf, prev_sumW, prev_M, prev_br, prev_bg, prev_bb,
                                             prev_cos);
                accr += hit.er * thr_r * w; accg += hit.eg * thr_g * w; accb += hit.eb * thr_b * w;
                adj.emission(hit, thr_r, thr_g, thr_b, w);
            }
            if (hit.ior > 1.0001f) {
                const DielectricSample ds =
                    sample_dielectric(wox, woy, woz, pnx, pny, pnz, dx, dy, dz, hit.ior, front, rough, alpha2);
                if (!ds.valid) break;
                adj.dielectric(hit, ds, accr, accg, accb);
                thr_r *= ds.eta_scale; thr_g *= ds.eta_scale; thr_b *= ds.eta_scale;
                float s = 1.0f;
                if (ds.n_x * pnx + (ds.n_y * pny + ds.n_z * pnz) < 0.0f) s = -1.0f;
                ox = hit.p.x + pnx * 0.001f * s;
                oy = hit.p.y + pny * 0.001f * s;
                oz = hit.p.z + pnz * 0.001f * s;
                dx = ds.n_x; dy = ds.n_y; dz = ds.n_z;
                prev_delta = 1;
                prev_sumW = 0.0f; prev_M = 0;
            } else {
                float spec_prob = 0.05f + 0.95f * metallic;
                if (spec_prob > 0.98f) spec_prob = 0.98f;
                if (metallic >= 0.999f) spec_prob = 1.0f;
                float cur_sumW = 0.0f;
                uint32_t cur_M = 0;
                {  // direct_lighting
                    const NeeSample ns = nee_prepare(hit.p.x, hit.p.y, hit.p.z, pnx, pny, pnz, wox, woy, woz, rough,
                                                     metallic, spec_prob, hit.ar, hit.ag, hit.ab);
                    const bool lit = adj.visible(*this, ns);
                    cur_sumW = ns.sumW;
                    cur_M = ns.M;
                    const float dlr = lit ? ns.r : 0.0f, dlg = lit ? ns.g : 0.0f, dlb = lit ? ns.b : 0.0f;
                    accr += dlr * thr_r; accg += dlg * thr_g; accb += dlb * thr_b;
                    if (lit) adj.nee(hit, ns, thr_r, thr_g, thr_b, wox, woy, woz, pnx, pny, pnz, alpha2, metallic);
                }
                float nx2, ny2, nz2;
                if (urand() < spec_prob) {
                    float hx, hy, hz;
                    sample_ggx_half(alpha2, pnx, pny, pnz, hx, hy, hz);
                    const float md = wox * hx + (woy * hy + woz * hz);
                    if (!(md > 1e-8f)) break;
                    nx2 = 2.0f * md * hx - wox; ny2 = 2.0f * md * hy - woy; nz2 = 2.0f * md * hz - woz;
                    prev_delta = (rough <= 0.0201f) ? 1 : 0;
                } else {
                    sample_cosine(pnx, pny, pnz, nx2, ny2, nz2);
                    prev_delta = 0;
                }
                const float ncos = nx2 * pnx + (ny2 * pny + nz2 * pnz);
                if (!(ncos > 1e-8f)) break;
                {
                    const float d = 1.0f / sqrtf(nx2 * nx2 + ny2 * ny2 + nz2 * nz2);
                    nx2 *= d; ny2 *= d; nz2 *= d;
                }
                const Bsdf e = eval_bsdf(nx2, ny2, nz2, wox, woy, woz, pnx, pny, pnz, alpha2, metallic, spec_prob,
                                         hit.ar, hit.ag, hit.ab);
                if (!(e.pdf > 1e-12f)) break;
                adj.scatter(hit, e, nx2, ny2, nz2, wox, woy, woz, pnx, pny, pnz, alpha2, metallic, accr, accg, accb);
                last_pdf = e.pdf;
                const float cos_i = nx2 * pnx + (ny2 * pny + nz2 * pnz);
                thr_r *= e.r * cos_i / e.pdf;
                thr_g *= e.g * cos_i / e.pdf;
                thr_b *= e.b * cos_i / e.pdf;
                prev_sumW = cur_sumW; prev_M = cur_M;
                prev_br = e.r; prev_bg = e.g; prev_bb = e.b; prev_cos = cos_i;
                ox = hit.p.x + pnx * 0.001f;
                oy = hit.p.y + pny * 0.001f;
                oz = hit.p.z + pnz * 0.001f;
                dx = nx2; dy = ny2; dz = nz2;
            }
            if (thr_r < 0.0f) thr_r = 0.0f;
            if (thr_g < 0.0f) thr_g = 0.0f;
            if (thr_b < 0.0f) thr_b = 0.0f;
            float mx = fmaxf(thr_r, thr_g);
            mx = fmaxf(mx, thr_b);
            if (!(mx >= 1e-12f)) break;
            if (bounce >= 1) {
                float rr = mx;
                if (rr < 0.05f) rr = 0.05f;
                if (rr > 0.9f) rr = 0.9f;
                if (!(urand() <= rr)) break;
                const float ir = 1.0f / rr;
                thr_r *= ir; thr_g *= ir; thr_b *= ir;
            }
            bounce++;
        }
        return {accr, accg, accb};
    }

    // Camera ray (sample_sensor, pathtrace.metal:729-736). generate selects the
    // (sx^2 + sy^2) + f^2 grouping of restir_generate (pathtrace.metal:834).
    __device__ V3 primary_dir(float sx, float sy, bool generate) const {
        const rm_camera& c = P.cam;
        const float f2 = c.focal * c.focal;
        const float d = 1.0f / sqrtf(generate ? (sx * sx + sy * sy) + f2 : sx * sx + (sy * sy + f2));
        const float tx = sx * d, ty = sy * d, tz = c.focal * d;
        return {tx * c.m0 + (ty * c.m1 + tz * c.m2), tx * c.m3 + (ty * c.m4 + tz * c.m5),
                tx * c.m6 + (ty * c.m7 + tz * c.m8)};
    }
    __device__ void jitter(uint32_t x, uint32_t y, float& sx, float& sy) {  // pathtrace.metal:754-755
        sx = float(x) + 1e-12f + (P.resolution - 2e-12f) * urand() - float(P.width) * 0.5f;
        sy = float(P.height) * 0.5f - (float(y) + 1e-12f + (P.resolution - 2e-12f) * urand());
    }
    __device__ V3 sample_sensor(float sx, float sy) {
        return pathtrace({P.cam.x, P.cam.y, P.cam.z}, primary_dir(sx, sy, false), P.bounces);
    }

    // Adds this lane's counts to the frame totals of its pass.
    __device__ void flush(unsigned long long* counters, int pass) {
        if (!counters) return;
        atomicAdd(counters + 3 * pass, (unsigned long long)tl.rays);
        atomicAdd(counters + 3 * pass + 1, (unsigned long long)tl.nodes);
        atomicAdd(counters + 3 * pass + 2, (unsigned long long)tl.tris);
        if (tl.overflow) atomicAdd(counters + kCntOverflow, 1ull);
    }
};

__device__ __forceinline__ uint8_t tonemap8(float c) {
    float v = floorf(c / (c + 1.0f) * 255.0f);
    v = fminf(fmaxf(v, 0.0f), 255.0f);
    return uint8_t(v);
}

}  // namespace rmh
