// GPU build of the 4-wide traversal tree from triangles. No counterpart in the
// reference, which builds a binary SAH tree on the host whenever the geometry
// topology changes and at every 32nd geometry update (its
// include/raymotion/runtime.hpp:119-130).
//
// Linear BVH: 63-bit Morton codes of the triangle centroids (21 bits per axis
// over the centroid bounds), a device radix sort of (code, triangle) pairs,
// the binary radix tree of the sorted codes built with one thread per inner
// node (Karras, "Maximizing Parallelism in the Construction of BVHs, Octrees,
// and k-d Trees", HPG 2012; equal codes ordered by position), bounds from the
// leaves up with one arrival counter per inner node, then a top-down collapse
// to 4-wide nodes, one launch per output level: each wide node opens the
// binary children of largest surface area until it has four, children of at
// most leaf_tris triangles become leaf slots and the others wide nodes of the
// next level. Nodes of one level are contiguous, which is the order the
// refit launches need. Traversal returns the closest hit whatever the tree,
// so images equal those rendered with a host-built tree.
#include <hipcub/hipcub.hpp>

#include "rm_kernels.h"

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

// Centroid bounds into cb[0..2] (min) and cb[3..5] (max), as ordered words.
__global__ __launch_bounds__(kB) void k_centroid_bounds(const float4* __restrict__ geo, int n, uint32_t* cb) {
    const int i = int(blockIdx.x) * kB + int(threadIdx.x);
    float lo[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
    float hi[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
    if (i < n) {
        float c[3];
        centroid(geo + size_t(i) * kTriGeoWords, c);
        for (int a = 0; a < 3; ++a) lo[a] = hi[a] = c[a];
    }
    for (int a = 0; a < 3; ++a) {
        lo[a] = wave_min(lo[a]);
        hi[a] = wave_max(hi[a]);
    }
    if (__lane_id() == 0 && i - int(__lane_id()) < n)
        for (int a = 0; a < 3; ++a) {
            atomicMin(cb + a, ordered(lo[a]));
            atomicMax(cb + 3 + a, ordered(hi[a]));
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

// Common-prefix length of sorted codes i and j (equal codes continue with the
// positions), -1 outside [0, n).
__device__ __forceinline__ int delta(const uint64_t* __restrict__ keys, int n, int i, long long j) {
    if (j < 0 || j >= n) return -1;
    const uint64_t a = keys[i], b = keys[j];
    if (a == b) return 64 + __clz(uint32_t(i) ^ uint32_t(j));
    return __clzll((long long)(a ^ b));
}

// Inner node i of the radix tree (n >= 2). Node ids: inner 0 .. n-2, leaf k
// as n - 1 + k. range[i] = first and last sorted triangle below.
__global__ __launch_bounds__(kB) void k_radix_tree(const uint64_t* __restrict__ keys, int n, int2* __restrict__ child,
                                                   int* __restrict__ parent, int2* __restrict__ range) {
    const int i = int(blockIdx.x) * kB + int(threadIdx.x);
    if (i >= n - 1) return;
    const int d = delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1) >= 0 ? 1 : -1;
    const int dmin = delta(keys, n, i, (long long)i - d);
    long long lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    long long l = 0;
    for (long long t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const long long j = i + l * d;
    const int dnode = delta(keys, n, i, j);
    long long s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int gamma = int(i + s * d + (d < 0 ? -1 : 0));
    const int first = int(min((long long)i, j)), last = int(max((long long)i, j));
    const int left = first == gamma ? n - 1 + gamma : gamma;
    const int right = last == gamma + 1 ? n - 1 + gamma + 1 : gamma + 1;
    child[i] = make_int2(left, right);
    range[i] = make_int2(first, last);
    parent[left] = i;
    parent[right] = i;
}

__device__ __forceinline__ void store_box(float* __restrict__ box, int node, const TriBox& b) {
    float* p = box + size_t(node) * 8;
    for (int a = 0; a < 3; ++a) {
        __hip_atomic_store(p + a, b.lo[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(p + 4 + a, b.hi[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}
__device__ __forceinline__ TriBox load_box(float* __restrict__ box, int node) {
    const float* p = box + size_t(node) * 8;
    TriBox b;
    for (int a = 0; a < 3; ++a) {
        b.lo[a] = __hip_atomic_load(p + a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        b.hi[a] = __hip_atomic_load(p + 4 + a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return b;
}

// Boxes of the inner nodes, from the leaves up: the second child to arrive
// at a node computes its box and continues to the parent.
__global__ __launch_bounds__(kB) void k_tree_bounds(const float4* __restrict__ geo, int n, const int2* __restrict__ child,
                                                    const int* __restrict__ parent, int* __restrict__ arrivals,
                                                    float* __restrict__ box) {
    const int k = int(blockIdx.x) * kB + int(threadIdx.x);
    if (k >= n) return;
    int node = parent[n - 1 + k];
    while (true) {
        __threadfence();
        if (atomicAdd(arrivals + node, 1) == 0) return;  // the other child finishes this node
        __threadfence();
        const int2 ch = child[node];
        const int c2[2] = {ch.x, ch.y};
        TriBox b;
        for (int a = 0; a < 3; ++a) {
            b.lo[a] = __builtin_inff();
            b.hi[a] = -__builtin_inff();
        }
        for (int c = 0; c < 2; ++c) {
            const TriBox cb = c2[c] >= n - 1 ? tri_box(geo + size_t(c2[c] - (n - 1)) * kTriGeoWords)
                                             : load_box(box, c2[c]);
            for (int a = 0; a < 3; ++a) {
                b.lo[a] = fminf(b.lo[a], cb.lo[a]);
                b.hi[a] = fmaxf(b.hi[a], cb.hi[a]);
            }
        }
        store_box(box, node, b);
        if (node == 0) return;
        node = parent[node];
    }
}

struct BuildTree {
    const float4* geo;  // leaf order
    int n;
    const int2* child;
    const int2* range;
    float* box;
};

__device__ __forceinline__ TriBox node_box(const BuildTree& T, int id) {
    return id >= T.n - 1 ? tri_box(T.geo + size_t(id - (T.n - 1)) * kTriGeoWords) : load_box(T.box, id);
}
__device__ __forceinline__ float half_area(const TriBox& b) {
    const float x = b.hi[0] - b.lo[0], y = b.hi[1] - b.lo[1], z = b.hi[2] - b.lo[2];
    return x * y + y * z + z * x;
}

// One output level: item i (binary node in[i].x, traversal stack depth above
// in[i].y) becomes wide node base + i. Inner children are appended to `out`
// (their wide index base + count + position) through *out_count.
__global__ __launch_bounds__(kB) void k_collapse(BuildTree T, const int2* __restrict__ in, int count, int base,
                                                 int2* __restrict__ out, unsigned* __restrict__ out_count,
                                                 float4* __restrict__ nodes, int* __restrict__ stack_need,
                                                 int leaf_tris) {
    const int i = int(blockIdx.x) * kB + int(threadIdx.x);
    if (i >= count) return;
    const int b = in[i].x, stack_here = in[i].y;
    int slot[4];
    int ns = 0;
    if (b >= T.n - 1) {
        slot[ns++] = b;  // a one-triangle tree
    } else {
        const int2 ch = T.child[b];
        slot[ns++] = ch.x;
        slot[ns++] = ch.y;
    }
    while (ns < 4) {
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
    float lo[3][4], hi[3][4];
    int ref[4], kind[4];
    int inner = 0;
    for (int s = 0; s < 4; ++s) {
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
        const int first = id >= T.n - 1 ? id - (T.n - 1) : T.range[id].x;
        const int tris = id >= T.n - 1 ? 1 : T.range[id].y - T.range[id].x + 1;
        if (tris <= leaf_tris) {
            ref[s] = first;
            kind[s] = tris;
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
    float4* w = nodes + size_t(base + i) * kNodeWords;
    for (int a = 0; a < 3; ++a) {
        w[2 * a] = make_float4(lo[a][0], lo[a][1], lo[a][2], lo[a][3]);
        w[2 * a + 1] = make_float4(hi[a][0], hi[a][1], hi[a][2], hi[a][3]);
    }
    w[6] = make_float4(__int_as_float(ref[0]), __int_as_float(ref[1]), __int_as_float(ref[2]), __int_as_float(ref[3]));
    w[7] = make_float4(__int_as_float(kind[0]), __int_as_float(kind[1]), __int_as_float(kind[2]),
                       __int_as_float(kind[3]));
}

unsigned blocks(size_t n) { return unsigned((n + kB - 1) / kB); }

}  // namespace

void launch_gather_tris(const float4* src, const uint32_t* perm, int n, float4* dst, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_gather, dim3(blocks(size_t(n) * kTriGeoWords)), dim3(kB), 0, st, src, perm, n, dst);
}

size_t gpu_build_bytes(int n) {
    const size_t N = size_t(n > 1 ? n : 2);
    size_t sort_bytes = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, static_cast<const uint64_t*>(nullptr),
                                             static_cast<uint64_t*>(nullptr), static_cast<const uint32_t*>(nullptr),
                                             static_cast<uint32_t*>(nullptr), int(N), 0, 63);
    // keys x2, values (the second is the permutation, kept), child, range,
    // parent, arrivals, boxes, two level lists, bounds, counters, sort scratch
    return N * (8 * 2 + 4 + 8 + 8 + 4 * 2 + 4 + 32 + 8 * 2) + 16 * 256 + ((sort_bytes + 255) & ~size_t(255));
}

int gpu_build(const GpuBuildArgs& a, GpuBuildResult& r, hipStream_t st) {
    const int n = a.n;
    r = GpuBuildResult{};
    if (n <= 0) return int(hipErrorInvalidValue);
    const size_t N = size_t(n > 1 ? n : 2);
    char* p = static_cast<char*>(a.scratch);
    auto take = [&](size_t bytes) {
        char* q = p;
        p += (bytes + 255) & ~size_t(255);
        return static_cast<void*>(q);
    };
    uint64_t* keys_in = static_cast<uint64_t*>(take(N * 8));
    uint64_t* keys = static_cast<uint64_t*>(take(N * 8));
    uint32_t* vals_in = static_cast<uint32_t*>(take(N * 4));
    int2* child = static_cast<int2*>(take(N * 8));
    int2* range = static_cast<int2*>(take(N * 8));
    int* parent = static_cast<int*>(take(N * 8));
    int* arrivals = static_cast<int*>(take(N * 4));
    float* box = static_cast<float*>(take(N * 32));
    int2* lvl[2] = {static_cast<int2*>(take(N * 8)), static_cast<int2*>(take(N * 8))};
    uint32_t* ctl = static_cast<uint32_t*>(take(64));  // 6 centroid-bound words, out count, stack need
    size_t sort_bytes = 0;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, keys_in, keys, vals_in, a.perm, n, 0, 63, st);
    if (e != hipSuccess) return int(e);
    void* sort_tmp = take(sort_bytes);
    if (size_t(p - static_cast<char*>(a.scratch)) > a.scratch_bytes) return int(hipErrorInvalidValue);

    const uint32_t init[8] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u, 0u, 0u};
    if ((e = hipMemcpyAsync(ctl, init, sizeof(init), hipMemcpyHostToDevice, st)) != hipSuccess) return int(e);
    hipLaunchKernelGGL(k_centroid_bounds, dim3(blocks(size_t(n))), dim3(kB), 0, st, a.geo_scene, n, ctl);
    hipLaunchKernelGGL(k_morton, dim3(blocks(size_t(n))), dim3(kB), 0, st, a.geo_scene, n, ctl, keys_in, vals_in);
    if ((e = hipcub::DeviceRadixSort::SortPairs(sort_tmp, sort_bytes, keys_in, keys, vals_in, a.perm, n, 0, 63, st)) !=
        hipSuccess)
        return int(e);
    hipLaunchKernelGGL(k_gather, dim3(blocks(size_t(n) * kTriGeoWords)), dim3(kB), 0, st, a.geo_scene, a.perm, n,
                       a.geo);
    if (n >= 2) {
        hipLaunchKernelGGL(k_radix_tree, dim3(blocks(size_t(n - 1))), dim3(kB), 0, st, keys, n, child, parent, range);
        if ((e = hipMemsetAsync(arrivals, 0, size_t(n) * 4, st)) != hipSuccess) return int(e);
        hipLaunchKernelGGL(k_tree_bounds, dim3(blocks(size_t(n))), dim3(kB), 0, st, a.geo, n, child, parent, arrivals,
                           box);
    }
    // Collapse, one level per launch; the host reads each level's size.
    const int2 root = make_int2(n >= 2 ? 0 : n - 1, 0);
    if ((e = hipMemcpyAsync(lvl[0], &root, sizeof(root), hipMemcpyHostToDevice, st)) != hipSuccess) return int(e);
    const BuildTree T{a.geo, n, child, range, box};
    int base = 0, count = 1, cur = 0;
    r.level_count.clear();
    while (count > 0) {
        if (size_t(base) + size_t(count) > a.max_nodes) return int(hipErrorInvalidValue);
        if ((e = hipMemsetAsync(ctl + 6, 0, 4, st)) != hipSuccess) return int(e);
        hipLaunchKernelGGL(k_collapse, dim3(blocks(size_t(count))), dim3(kB), 0, st, T, lvl[cur], count, base,
                           lvl[cur ^ 1], ctl + 6, a.nodes, reinterpret_cast<int*>(ctl + 7), a.leaf_tris);
        if ((e = hipMemcpyAsync(a.host_words, ctl + 6, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, st)) !=
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
    return int(hipGetLastError());
}

}  // namespace rmh
