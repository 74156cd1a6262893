// GPU build of the wide traversal tree from triangles. No counterpart in the
// reference, which builds a binary SAH tree on the host whenever the geometry
// topology changes and at every 32nd geometry update (its
// include/raymotion/runtime.hpp:119-130).
//
// 63-bit Morton codes of the triangle centroids (21 bits per axis over the
// centroid bounds) and a device radix sort (rm_sort.hip) put the triangles in
// spatial order; the binary tree is then built bottom-up by parallel
// locally-ordered clustering (Meister and Bittner, "Parallel Locally-Ordered Clustering for
// Bounding Volume Hierarchy Construction", IEEE TVCG 2018): every cluster
// finds, within kPlocRadius neighbours on either side in that order, the one
// whose merged box has the least surface area; mutual nearest neighbours merge
// into a new node, and the surviving clusters are compacted in order, until
// one is left. A top-down collapse then writes kWide-wide nodes, one launch per
// output level: each wide node opens the binary children of largest surface
// area until all kWide slots are used; single triangles become leaf slots, the others
// nodes of the next level. The nodes are then renumbered depth-first (a node
// next to its first child, as in the host tree's preorder) and listed level by
// level for the refit launches. Traversal returns the closest hit whatever the
// tree, so images equal those rendered with a host-built tree.
#include <algorithm>

#include "rm_kernels.h"
#include "rm_sort.h"

namespace rmh {
namespace {

constexpr int kB = 256;

__device__ __forceinline__ uint32_t ordered(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unordered(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

struct TriBox {
    float lo[3], hi[3];
};

__device__ __forceinline__ TriBox tri_box(const float4* __restrict__ g) {
    const float4 g0 = g[0], g1 = g[1], g2 = g[2];
    TriBox b;
    b.lo[0] = fminf(fminf(g0.x, g0.w), g1.z); b.hi[0] = fmaxf(fmaxf(g0.x, g0.w), g1.z);
    b.lo[1] = fminf(fminf(g0.y, g1.x), g1.w); b.hi[1] = fmaxf(fmaxf(g0.y, g1.x), g1.w);
    b.lo[2] = fminf(fminf(g0.z, g1.y), g2.x); b.hi[2] = fmaxf(fmaxf(g0.z, g1.y), g2.x);
    return b;
}

// Unscaled centroid v0 + v1 + v2 (the scale cancels in the quantisation).
__device__ __forceinline__ void centroid(const float4* __restrict__ g, float c[3]) {
    const float4 g0 = g[0], g1 = g[1], g2 = g[2];
    c[0] = (g0.x + g0.w) + g1.z;
    c[1] = (g0.y + g1.x) + g1.w;
    c[2] = (g0.z + g1.y) + g2.x;
}

__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

// Centroid bounds into cb[0..2] (min) and cb[3..5] (max), as ordered words:
// a grid-stride pass, wave then block reductions, six atomics per block.
__global__ __launch_bounds__(kB) void k_centroid_bounds(const float4* __restrict__ geo, int n, uint32_t* cb) {
    __shared__ float red[kB / 64][6];
    float lo[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
    float hi[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
    for (int i = int(blockIdx.x) * kB + int(threadIdx.x); i < n; i += int(gridDim.x) * kB) {
        float c[3];
        centroid(geo + size_t(i) * kTriGeoWords, c);
        for (int a = 0; a < 3; ++a) {
            lo[a] = fminf(lo[a], c[a]);
            hi[a] = fmaxf(hi[a], c[a]);
        }
    }
    for (int a = 0; a < 3; ++a) {
        lo[a] = wave_min(lo[a]);
        hi[a] = wave_max(hi[a]);
    }
    const int wave = int(threadIdx.x) / 64;
    if (__lane_id() == 0)
        for (int a = 0; a < 3; ++a) {
            red[wave][a] = lo[a];
            red[wave][3 + a] = hi[a];
        }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = int(threadIdx.x);
        float v = red[0][k];
        for (int w = 1; w < kB / 64; ++w) v = k < 3 ? fminf(v, red[w][k]) : fmaxf(v, red[w][k]);
        if (v == v && (k < 3 ? v < __builtin_inff() : v > -__builtin_inff())) {  // blocks without triangles skip
            if (k < 3) atomicMin(cb + k, ordered(v));
            else atomicMax(cb + k, ordered(v));
        }
    }
}

__device__ __forceinline__ uint64_t spread21(uint32_t v) {
    uint64_t x = v & 0x1FFFFFu;
    x = (x | (x << 32)) & 0x1F00000000FFFFULL;
    x = (x | (x << 16)) & 0x1F0000FF0000FFULL;
    x = (x | (x << 8)) & 0x100F00F00F00F00FULL;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ULL;
    x = (x | (x << 2)) & 0x1249249249249249ULL;
    return x;
}

__global__ __launch_bounds__(kB) void k_morton(const float4* __restrict__ geo, int n, const uint32_t* __restrict__ cb,
                                               uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const int i = int(blockIdx.x) * kB + int(threadIdx.x);
    if (i >= n) return;
    float c[3];
    centroid(geo + size_t(i) * kTriGeoWords, c);
    uint32_t q[3];
    for (int a = 0; a < 3; ++a) {
        const float lo = unordered(cb[a]), hi = unordered(cb[3 + a]);
        float t = hi > lo ? (c[a] - lo) / (hi - lo) : 0.0f;
        t = fminf(fmaxf(t, 0.0f), 1.0f);
        q[a] = min(uint32_t(t * 2097152.0f), 2097151u);
    }
    keys[i] = (spread21(q[0]) << 2) | (spread21(q[1]) << 1) | spread21(q[2]);
    vals[i] = uint32_t(i);
}

// Triangle records in sorted (leaf) order.
__global__ __launch_bounds__(kB) void k_gather(const float4* __restrict__ src, const uint32_t* __restrict__ perm,
                                               int n, float4* __restrict__ dst) {
    const int w = int(blockIdx.x) * kB + int(threadIdx.x);
    if (w >= n * kTriGeoWords) return;
    const int k = w / kTriGeoWords, part = w - k * kTriGeoWords;
    dst[w] = src[size_t(perm[k]) * kTriGeoWords + part];
}

__device__ __forceinline__ void store_box(float* __restrict__ box, int node, const TriBox& b) {
    float4* p = reinterpret_cast<float4*>(box + size_t(node) * 8);
    p[0] = make_float4(b.lo[0], b.lo[1], b.lo[2], 0.0f);
    p[1] = make_float4(b.hi[0], b.hi[1], b.hi[2], 0.0f);
}
__device__ __forceinline__ TriBox load_box(const float* __restrict__ box, int node) {
    const float4* p = reinterpret_cast<const float4*>(box + size_t(node) * 8);
    const float4 l = p[0], h = p[1];
    TriBox b;
    b.lo[0] = l.x; b.lo[1] = l.y; b.lo[2] = l.z;
    b.hi[0] = h.x; b.hi[1] = h.y; b.hi[2] = h.z;
    return b;
}
__device__ __forceinline__ float half_area(const TriBox& b) {
    const float x = b.hi[0] - b.lo[0], y = b.hi[1] - b.lo[1], z = b.hi[2] - b.lo[2];
    return x * y + y * z + z * x;
}

constexpr int kPlocRadius = 8;  // measured against 16 and 32: equal or faster frames, faster build

// Leaf clusters in sorted order: id n - 1 + k for sorted triangle k (inner
// nodes take ids 0 .. n - 2 in order of creation).
__global__ __launch_bounds__(kB) void k_ploc_init(const float4* __restrict__ geo, int n, int* __restrict__ cid,
                                                  float4* __restrict__ clo, float4* __restrict__ chi) {
    const int k = int(blockIdx.x) * kB + int(threadIdx.x);
    if (k >= n) return;
    const TriBox b = tri_box(geo + size_t(k) * kTriGeoWords);
    cid[k] = n - 1 + k;
    clo[k] = make_float4(b.lo[0], b.lo[1], b.lo[2], 0.0f);
    chi[k] = make_float4(b.hi[0], b.hi[1], b.hi[2], 0.0f);
}

// Nearest neighbour of every cluster within kPlocRadius positions: least
// merged surface area, ties to the pair of smaller (lower, then higher) index,
// a total order on pairs, so the best pair overall is always mutual.
__global__ __launch_bounds__(kB) void k_ploc_nn(const float4* __restrict__ clo, const float4* __restrict__ chi,
                                                const unsigned* __restrict__ count, int* __restrict__ nn) {
    __shared__ float4 slo[kB + 2 * kPlocRadius], shi[kB + 2 * kPlocRadius];
    const int m = int(*count);
    const int b0 = int(blockIdx.x) * kB;
    if (b0 >= m) return;
    for (int t = int(threadIdx.x); t < kB + 2 * kPlocRadius; t += kB) {
        const int g = b0 - kPlocRadius + t;
        if (g >= 0 && g < m) {
            slo[t] = clo[g];
            shi[t] = chi[g];
        }
    }
    __syncthreads();
    const int i = b0 + int(threadIdx.x);
    if (i >= m) return;
    const int li = int(threadIdx.x) + kPlocRadius;
    const float4 al = slo[li], ah = shi[li];
    float best = __builtin_inff();
    int bj = -1;
    const int j0 = max(0, i - kPlocRadius), j1 = min(m - 1, i + kPlocRadius);
    for (int j = j0; j <= j1; ++j) {
        if (j == i) continue;
        const float4 bl = slo[j - b0 + kPlocRadius], bh = shi[j - b0 + kPlocRadius];
        const float x = fmaxf(ah.x, bh.x) - fminf(al.x, bl.x);
        const float y = fmaxf(ah.y, bh.y) - fminf(al.y, bl.y);
        const float z = fmaxf(ah.z, bh.z) - fminf(al.z, bl.z);
        const float c = x * y + y * z + z * x;
        // Candidates come in increasing j: for equal cost the first one has the
        // smaller pair (min(i, j), max(i, j)) as well.
        if (c < best) {
            best = c;
            bj = j;
        }
    }
    nn[i] = bj;
}

// Mutual pairs (i < j) merge into a new inner node held by cluster i; cluster
// j is dropped (keep = 0).
__global__ __launch_bounds__(kB) void k_ploc_merge(const unsigned* __restrict__ count, int* __restrict__ cid,
                                                   float4* __restrict__ clo, float4* __restrict__ chi,
                                                   const int* __restrict__ nn, int* __restrict__ keep,
                                                   unsigned* __restrict__ next_id, int2* __restrict__ child,
                                                   float* __restrict__ box) {
    const int m = int(*count);
    const int i = int(blockIdx.x) * kB + int(threadIdx.x);
    if (i >= m) return;
    const int j = nn[i];
    const bool mutual = j >= 0 && nn[j] == i;
    if (!mutual) {
        keep[i] = 1;
        return;
    }
    if (i > j) {
        keep[i] = 0;
        return;
    }
    const int a = cid[i], b = cid[j];
    const int node = int(atomicAdd(next_id, 1u));
    const float4 al = clo[i], ah = chi[i], bl = clo[j], bh = chi[j];
    TriBox u;
    u.lo[0] = fminf(al.x, bl.x); u.lo[1] = fminf(al.y, bl.y); u.lo[2] = fminf(al.z, bl.z);
    u.hi[0] = fmaxf(ah.x, bh.x); u.hi[1] = fmaxf(ah.y, bh.y); u.hi[2] = fmaxf(ah.z, bh.z);
    child[node] = make_int2(a, b);
    store_box(box, node, u);
    cid[i] = node;
    clo[i] = make_float4(u.lo[0], u.lo[1], u.lo[2], 0.0f);
    chi[i] = make_float4(u.hi[0], u.hi[1], u.hi[2], 0.0f);
    keep[i] = 1;
}

// Surviving clusters to their positions (exclusive sum of keep), order kept;
// the last thread writes the new count.
__global__ __launch_bounds__(kB) void k_ploc_compact(const unsigned* __restrict__ count, const int* __restrict__ keep,
                                                     const int* __restrict__ pos, const int* __restrict__ cid,
                                                     const float4* __restrict__ clo, const float4* __restrict__ chi,
                                                     int* __restrict__ cid2, float4* __restrict__ clo2,
                                                     float4* __restrict__ chi2, unsigned* __restrict__ count2) {
    const int m = int(*count);
    const int i = int(blockIdx.x) * kB + int(threadIdx.x);
    if (i >= m) return;
    if (keep[i]) {
        const int p = pos[i];
        cid2[p] = cid[i];
        clo2[p] = clo[i];
        chi2[p] = chi[i];
    }
    if (i == m - 1) *count2 = unsigned(pos[i] + keep[i]);
}

struct BuildTree {
    const float4* geo;  // leaf order
    int n;
    const int2* child;
    const float* box;
};

__device__ __forceinline__ TriBox node_box(const BuildTree& T, int id) {
    return id >= T.n - 1 ? tri_box(T.geo + size_t(id - (T.n - 1)) * kTriGeoWords) : load_box(T.box, id);
}

// One output level: item i (binary node in[i].x, traversal stack depth above
// in[i].y) becomes wide node base + i. Inner children are appended to `out`
// (their wide index base + count + position) through *out_count.
__global__ __launch_bounds__(kB) void k_collapse(BuildTree T, const int2* __restrict__ in, int count, int base,
                                                 int2* __restrict__ out, unsigned* __restrict__ out_count,
                                                 float4* __restrict__ nodes, int* __restrict__ stack_need) {
    const int i = int(blockIdx.x) * kB + int(threadIdx.x);
    if (i >= count) return;
    const int b = in[i].x, stack_here = in[i].y;
    int slot[kWide];
    int ns = 0;
    if (b >= T.n - 1) {
        slot[ns++] = b;  // a one-triangle tree
    } else {
        const int2 ch = T.child[b];
        slot[ns++] = ch.x;
        slot[ns++] = ch.y;
    }
    while (ns < kWide) {
        int best = -1;
        float best_area = -1.0f;
        for (int s = 0; s < ns; ++s) {
            if (slot[s] >= T.n - 1) continue;  // a single triangle
            const float a = half_area(load_box(T.box, slot[s]));
            if (a > best_area) {
                best_area = a;
                best = s;
            }
        }
        if (best < 0) break;
        const int2 ch = T.child[slot[best]];
        slot[best] = ch.x;
        slot[ns++] = ch.y;
    }
    float lo[3][kWide], hi[3][kWide];
    int ref[kWide], kind[kWide];
    int inner = 0;
    for (int s = 0; s < kWide; ++s) {
        if (s >= ns) {
            for (int a = 0; a < 3; ++a) {
                lo[a][s] = __builtin_inff();
                hi[a][s] = -__builtin_inff();
            }
            ref[s] = 0;
            kind[s] = -1;
            continue;
        }
        const int id = slot[s];
        const TriBox bx = node_box(T, id);
        for (int a = 0; a < 3; ++a) {
            lo[a][s] = bx.lo[a];
            hi[a][s] = bx.hi[a];
        }
        if (id >= T.n - 1) {
            ref[s] = id - (T.n - 1);  // leaf-order triangle
            kind[s] = 1;
        } else {
            ++inner;
            kind[s] = 0;
            ref[s] = -1;  // below
        }
    }
    const int stack_below = stack_here + max(0, inner - 1);
    if (inner > 0) {
        atomicMax(stack_need, stack_below);
        const unsigned pos = atomicAdd(out_count, unsigned(inner));
        int k = 0;
        for (int s = 0; s < ns; ++s)
            if (kind[s] == 0) {
                out[pos + k] = make_int2(slot[s], stack_below);
                ref[s] = base + count + int(pos) + k;
                ++k;
            }
    }
    float4 w[kNodeWords];
    write_node(w, lo, hi, ref, kind);
    float4* dst = nodes + size_t(base + i) * kNodeWords;
    for (int k = 0; k < kNodeWords; ++k) dst[k] = w[k];
}

// Depth-first renumbering of the collapsed tree (which comes out level by
// level): subtree sizes from the deepest level up, then indices from the root
// down, each inner child's subtree following its parent's earlier children, so
// a node sits next to its first child as in the host tree's preorder.
__global__ __launch_bounds__(kB) void k_subtree_size(const float4* __restrict__ nodes, int first, int count,
                                                     int* __restrict__ size) {
    const int i = int(blockIdx.x) * kB + int(threadIdx.x);
    if (i >= count) return;
    int ref[kWide], kind[kWide];
    read_node_links(nodes + size_t(first + i) * kNodeWords, ref, kind);
    int s = 1;
    for (int c = 0; c < kWide; ++c)
        if (kind[c] == 0) s += size[ref[c]];
    size[first + i] = s;
}

__global__ __launch_bounds__(kB) void k_dfs_index(const float4* __restrict__ nodes, int first, int count,
                                                  const int* __restrict__ size, int* __restrict__ index) {
    const int i = int(blockIdx.x) * kB + int(threadIdx.x);
    if (i >= count) return;
    int ref[kWide], kind[kWide];
    read_node_links(nodes + size_t(first + i) * kNodeWords, ref, kind);
    int next = index[first + i] + 1;
    for (int c = 0; c < kWide; ++c)
        if (kind[c] == 0) {
            index[ref[c]] = next;
            next += size[ref[c]];
        }
}

// Node i to index[i], inner-child references renumbered.
__global__ __launch_bounds__(kB) void k_renumber(const float4* __restrict__ src, int total,
                                                 const int* __restrict__ index, float4* __restrict__ dst) {
    const int i = int(blockIdx.x) * kB + int(threadIdx.x);
    if (i >= total) return;
    float4 w[kNodeWords];
    const float4* a = src + size_t(i) * kNodeWords;
    for (int k = 0; k < kNodeWords; ++k) w[k] = a[k];
    int ref[kWide], kind[kWide];
    read_node_links(w, ref, kind);
    for (int c = 0; c < kWide; ++c)
        if (kind[c] == 0) ref[c] = index[ref[c]];
    set_node_refs(w, ref);
    float4* b = dst + size_t(index[i]) * kNodeWords;
    for (int k = 0; k < kNodeWords; ++k) b[k] = w[k];
}

// Refit order: the new indices of one level's nodes at out[0 .. count).
__global__ __launch_bounds__(kB) void k_level_order(const int* __restrict__ index, int first, int count,
                                                    int32_t* __restrict__ out) {
    const int i = int(blockIdx.x) * kB + int(threadIdx.x);
    if (i < count) out[i] = index[first + i];
}

unsigned blocks(size_t n) { return unsigned((n + kB - 1) / kB); }

// Device scratch layout of one build (256-byte aligned pieces).
struct Scratch {
    uint64_t *keys_in, *keys;
    uint32_t* vals_in;
    int2* child;
    float* box;
    int* cid[2];
    float4 *clo[2], *chi[2];
    int *nn, *keep, *pos;
    int2* lvl[2];
    float4* wide;        // collapsed nodes in level order, before renumbering
    int *size, *index;   // per collapsed node
    uint32_t* ctl;  // 6 centroid-bound words, cluster counts (2), next inner id, level count, stack need
    void *sort_tmp, *scan_tmp;
    size_t sort_bytes, scan_bytes, total;
};

Scratch carve(void* base, int n) {
    const size_t N = size_t(n > 1 ? n : 2);  // also bounds the wide node count (n - 1, or 1)
    Scratch s{};
    s.sort_bytes = sort_tmp_bytes(int(N));
    s.scan_bytes = scan_tmp_bytes(int(N));
    size_t off = 0;  // with base == nullptr only the size is wanted
    auto take = [&](size_t bytes) {
        void* q = base ? static_cast<void*>(static_cast<char*>(base) + off) : nullptr;
        off += (bytes + 255) & ~size_t(255);
        return q;
    };
    s.keys_in = static_cast<uint64_t*>(take(N * 8));
    s.keys = static_cast<uint64_t*>(take(N * 8));
    s.vals_in = static_cast<uint32_t*>(take(N * 4));
    s.child = static_cast<int2*>(take(N * 8));
    s.box = static_cast<float*>(take(N * 32));
    for (int k = 0; k < 2; ++k) {
        s.cid[k] = static_cast<int*>(take(N * 4));
        s.clo[k] = static_cast<float4*>(take(N * 16));
        s.chi[k] = static_cast<float4*>(take(N * 16));
        s.lvl[k] = static_cast<int2*>(take(N * 8));
    }
    s.nn = static_cast<int*>(take(N * 4));
    s.keep = static_cast<int*>(take(N * 4));
    s.pos = static_cast<int*>(take(N * 4));
    s.wide = static_cast<float4*>(take(N * kNodeWords * sizeof(float4)));
    s.size = static_cast<int*>(take(N * 4));
    s.index = static_cast<int*>(take(N * 4));
    s.ctl = static_cast<uint32_t*>(take(64));
    s.sort_tmp = take(s.sort_bytes);
    s.scan_tmp = take(s.scan_bytes);
    s.total = off;
    return s;
}

enum : int { kCtlCount0 = 6, kCtlCount1 = 7, kCtlNextId = 8, kCtlLevel = 9, kCtlStack = 10 };

}  // namespace

void launch_gather_tris(const float4* src, const uint32_t* perm, int n, float4* dst, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_gather, dim3(blocks(size_t(n) * kTriGeoWords)), dim3(kB), 0, st, src, perm, n, dst);
}

size_t gpu_build_bytes(int n) { return carve(nullptr, n).total; }

int gpu_build(const GpuBuildArgs& a, GpuBuildResult& r, hipStream_t st) {
    const int n = a.n;
    r = GpuBuildResult{};
    if (n <= 0) return int(hipErrorInvalidValue);
    Scratch s = carve(a.scratch, n);
    if (s.total > a.scratch_bytes) return int(hipErrorInvalidValue);
    hipError_t e;
    uint32_t init[11] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u, uint32_t(n), 0u, 0u, 0u, 0u};
    if ((e = hipMemcpyAsync(s.ctl, init, sizeof(init), hipMemcpyHostToDevice, st)) != hipSuccess) return int(e);
    const unsigned bound_blocks = std::min(blocks(size_t(n)), 1024u);
    hipLaunchKernelGGL(k_centroid_bounds, dim3(bound_blocks), dim3(kB), 0, st, a.geo_scene, n, s.ctl);
    hipLaunchKernelGGL(k_morton, dim3(blocks(size_t(n))), dim3(kB), 0, st, a.geo_scene, n, s.ctl, s.keys_in,
                       s.vals_in);
    if ((e = sort_pairs_63(s.keys_in, s.vals_in, s.keys, a.perm, n, s.sort_tmp, st)) != hipSuccess) return int(e);
    hipLaunchKernelGGL(k_gather, dim3(blocks(size_t(n) * kTriGeoWords)), dim3(kB), 0, st, a.geo_scene, a.perm, n,
                       a.geo);
    // Clustering. The host reads the cluster count after every round (it
    // only sizes the grids and ends the loop; the kernels read it on the GPU).
    // The collapse starts from the root: the last cluster, or the only leaf.
    const int2 leaf_root = make_int2(0, 0);
    if ((e = hipMemcpyAsync(s.lvl[0], &leaf_root, sizeof(leaf_root), hipMemcpyHostToDevice, st)) != hipSuccess)
        return int(e);
    if (n >= 2) {
        hipLaunchKernelGGL(k_ploc_init, dim3(blocks(size_t(n))), dim3(kB), 0, st, a.geo, n, s.cid[0], s.clo[0],
                           s.chi[0]);
        int m = n, cur = 0, rounds = 0;
        while (m > 1) {
            if (++rounds > 100000) return int(hipErrorUnknown);
            unsigned* cnt = s.ctl + kCtlCount0 + cur;
            unsigned* cnt2 = s.ctl + kCtlCount0 + (cur ^ 1);
            const dim3 g(blocks(size_t(m)));
            hipLaunchKernelGGL(k_ploc_nn, g, dim3(kB), 0, st, s.clo[cur], s.chi[cur], cnt, s.nn);
            hipLaunchKernelGGL(k_ploc_merge, g, dim3(kB), 0, st, cnt, s.cid[cur], s.clo[cur], s.chi[cur], s.nn,
                               s.keep, s.ctl + kCtlNextId, s.child, s.box);
            if ((e = exclusive_scan(reinterpret_cast<const uint32_t*>(s.keep), reinterpret_cast<uint32_t*>(s.pos), m,
                                    s.scan_tmp, st)) != hipSuccess)
                return int(e);
            hipLaunchKernelGGL(k_ploc_compact, g, dim3(kB), 0, st, cnt, s.keep, s.pos, s.cid[cur], s.clo[cur],
                               s.chi[cur], s.cid[cur ^ 1], s.clo[cur ^ 1], s.chi[cur ^ 1], cnt2);
            if ((e = hipMemcpyAsync(a.host_words, cnt2, sizeof(uint32_t), hipMemcpyDeviceToHost, st)) != hipSuccess)
                return int(e);
            if ((e = hipStreamSynchronize(st)) != hipSuccess) return int(e);
            const int m2 = int(a.host_words[0]);
            if (m2 >= m || m2 < 1) return int(hipErrorUnknown);  // every round merges at least one pair
            m = m2;
            cur ^= 1;
        }
        if ((e = hipMemcpyAsync(s.lvl[0], s.cid[cur], sizeof(int), hipMemcpyDeviceToDevice, st)) != hipSuccess)
            return int(e);
    }
    // Collapse, one level per launch; the host reads each level's size.
    const BuildTree T{a.geo, n, s.child, s.box};
    int base = 0, count = 1, cur = 0;
    r.level_count.clear();
    while (count > 0) {
        if (size_t(base) + size_t(count) > a.max_nodes) return int(hipErrorInvalidValue);
        if ((e = hipMemsetAsync(s.ctl + kCtlLevel, 0, 4, st)) != hipSuccess) return int(e);
        hipLaunchKernelGGL(k_collapse, dim3(blocks(size_t(count))), dim3(kB), 0, st, T, s.lvl[cur], count, base,
                           s.lvl[cur ^ 1], s.ctl + kCtlLevel, s.wide, reinterpret_cast<int*>(s.ctl + kCtlStack));
        if ((e = hipMemcpyAsync(a.host_words, s.ctl + kCtlLevel, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, st)) !=
            hipSuccess)
            return int(e);
        if ((e = hipStreamSynchronize(st)) != hipSuccess) return int(e);
        r.level_count.push_back(count);
        base += count;
        count = int(a.host_words[0]);
        cur ^= 1;
    }
    r.nodes = base;
    r.stack_need = int(a.host_words[1]);
    // Depth-first renumbering, and the refit order (deepest level first).
    const int levels = int(r.level_count.size());
    std::vector<int> first(size_t(levels), 0);
    for (int l = 1; l < levels; ++l) first[size_t(l)] = first[size_t(l - 1)] + r.level_count[size_t(l - 1)];
    for (int l = levels; l-- > 0;)
        hipLaunchKernelGGL(k_subtree_size, dim3(blocks(size_t(r.level_count[size_t(l)]))), dim3(kB), 0, st, s.wide,
                           first[size_t(l)], r.level_count[size_t(l)], s.size);
    if ((e = hipMemsetAsync(s.index, 0, sizeof(int), st)) != hipSuccess) return int(e);  // the root stays node 0
    for (int l = 0; l < levels; ++l)
        hipLaunchKernelGGL(k_dfs_index, dim3(blocks(size_t(r.level_count[size_t(l)]))), dim3(kB), 0, st, s.wide,
                           first[size_t(l)], r.level_count[size_t(l)], s.size, s.index);
    hipLaunchKernelGGL(k_renumber, dim3(blocks(size_t(r.nodes))), dim3(kB), 0, st, s.wide, r.nodes, s.index, a.nodes);
    int at = 0;
    for (int l = levels; l-- > 0;) {
        hipLaunchKernelGGL(k_level_order, dim3(blocks(size_t(r.level_count[size_t(l)]))), dim3(kB), 0, st, s.index,
                           first[size_t(l)], r.level_count[size_t(l)], a.level_nodes + at);
        at += r.level_count[size_t(l)];
    }
    return int(hipGetLastError());
}

}  // namespace rmh
