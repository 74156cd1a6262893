// Host-side launchers of the gfx950 kernels (rm_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>
#include <vector>

#include "rm_layout.h"

namespace rmh {

// `stack` is the traversal stack depth the resident tree needs; it selects the
// kernel instance with a large enough stack: 32, 48 or 64 entries per lane in
// LDS, or for deeper trees kSpillStack entries of which those past kLdsStack
// live in per-lane scratch memory. (A 48-entry LDS stack with deeper entries
// in scratch was measured 5-7 % slower on restir_shade, although rays rarely
// reach it, so trees that fit in LDS keep the LDS-only instances.) An
// instance holds at least one entry more than the depth: the traversal stores
// into the entry above the live ones on every step (Tracer::trav_node).
constexpr int kMaxStack = kSpillStack - 1;

// Calls f(std::integral_constant<int, S>{}) with the smallest stack capacity
// S in {32, 48, 64, kSpillStack} that exceeds `stack`.
template <class F>
inline auto with_stack(int stack, F&& f) {
    if (stack < 32) return f(std::integral_constant<int, 32>{});
    if (stack < 48) return f(std::integral_constant<int, 48>{});
    if (stack < kLdsStack) return f(std::integral_constant<int, kLdsStack>{});
    return f(std::integral_constant<int, kSpillStack>{});
}

// Work queues of the persistent passes: kBands bands of image rows, one per
// XCD (workgroups are dispatched to the 8 XCDs in turn). The queue buffer holds
// kQueueWords words, zero when a frame starts: band counters of generate, of
// shade / render_paths, and restir_shade's two order counts per band.
constexpr int kBands = 8;
constexpr int kQueueGenerate = 0, kQueuePath = kBands, kQueueOrder = 2 * kBands, kQueueWords = 4 * kBands;

// render_paths_simd with 8 or 32 lanes per pixel.
void launch_render_paths(int stack, int lanes, const DevScene& sc, const FrameParams& P, uint8_t* out, float4* hdr,
                         unsigned long long* cnt, hipStream_t st);
// Persistent restir_generate; `queue` (kBands words) must hold 0 when the kernel starts.
void launch_generate(int stack, const DevScene& sc, const FrameParams& P, const GBufferSoA& gb, const ReservoirSoA& ri,
                     unsigned int* queue, unsigned long long* cnt, hipStream_t st);
void launch_spatial(int stack, const DevScene& sc, const FrameParams& P, const GBufferSoA& gb, const ReservoirSoA& ri,
                    const ReservoirSoA& ro, const GBufferSoA& hg, const ReservoirSoA& hr, unsigned long long* cnt,
                    hipStream_t st);
// Persistent path engine: restir_shade (shade = true, writes hdr and, if not
// null, the per-pixel sample variance of luminance) or render_paths with one
// lane per pixel (writes rgb, and hdr if not null). `queue` (kBands words) and,
// for restir_shade with `order` (width * height words), `order_n` (2 * kBands
// words) must hold 0 when the pass starts; restir_shade then queues each
// band's pixels by expected chain length.
void launch_wave(int stack, bool shade, const DevScene& sc, const FrameParams& P, const GBufferSoA& gb,
                 const ReservoirSoA& rm, float4* hdr, uint8_t* rgb, float* lum_var, unsigned int* queue,
                 unsigned int* order, unsigned int* order_n, unsigned long long* cnt, hipStream_t st);
// Wavefront variant of launch_wave: the same per-pixel results from a lean
// trace kernel and a state-machine kernel over path states kept in HBM.
struct WaveBufs {
    uint32_t* state[2];  // wave_slot_words(shade) * stride words each
    uint4* hits;         // stride entries
    unsigned int* ctl;   // 4 control words
    unsigned int* host;  // 2 words of pinned host memory
    hipEvent_t ev[2];
    size_t stride;       // slot capacity, at least width * height
    bool tail;           // hand the last slots to the persistent loop (else iterate to the end)
};
struct WaveRun {
    int iters;      // trace + step iterations launched
    unsigned tail;  // active slots (last read-back) handed to the persistent loop
};
size_t wave_slot_words(bool shade);
// Returns a hipError_t code; waits on the stream a few times.
int run_wave(int stack, bool shade, const DevScene& sc, const FrameParams& P, const GBufferSoA& gb,
             const ReservoirSoA& rm, float4* hdr, uint8_t* rgb, float* lum_var, const WaveBufs& wb,
             unsigned long long* cnt, hipStream_t st, WaveRun* run);
void launch_temporal(const FrameParams& P, const GBufferSoA& gb, const GBufferSoA& hg, const float4* in,
                     const float4* hist, float4* out, hipStream_t st);
void launch_atrous(const FrameParams& P, const GBufferSoA& gb, const float4* in, float4* out, int step, hipStream_t st);
void launch_tonemap(int npx, const float4* in, uint8_t* rgb, float* hdr3, hipStream_t st);
void launch_zero_u64(unsigned long long* p, int n, hipStream_t st);
void launch_trace(int stack, const DevScene& sc, const FrameParams& P, int n, const float* o, const float* d,
                  const float* tmax, int any, int32_t* tri, float4* tuv, unsigned long long* cnt, hipStream_t st);
// Path-replay adjoint of render_paths (1, 8 or 32 lanes per pixel, no adaptive
// stop): gradients of sum(d_img * image) with respect to per-group material
// parameters, accumulated into grad (5 doubles per group, zeroed by the caller).
constexpr int kMaxGroups = 128;
struct AdjointArgs {
    const float* d_img;          // 3 per pixel: dL/d(pixel radiance)
    const int32_t* tri_group;    // per triangle in scene order; -1: no parameters
    const int32_t* light_group;  // per light in CDF order
    const float* group_base;     // 3 per group: emission = scale * base
    int ngroup;
    double* grad;                // albedo r, g, b, roughness, emission scale per group
    unsigned long long* mismatch;  // samples whose replay differed from their primal value
    // One record per (pixel, lane) stream for the sample round in flight:
    // the primal radiance (xyz), the RNG state at the sample's start and
    // after it, and the shadow-test outcomes of its path (two words).
    float4* primal;
    uint4* record;               // start state, next state, outcome bits 0-31, 32-63
};
void launch_adjoint(int stack, int lanes, const DevScene& sc, const FrameParams& P, const AdjointArgs& A,
                    unsigned long long* cnt, hipStream_t st);
// GPU tree build (rm_build.hip): triangles in Morton order, clustered into a
// binary tree, collapsed to 4-wide nodes with one triangle per leaf slot.
struct GpuBuildArgs {
    const float4* geo_scene;  // n triangle records, scene order (input)
    float4* geo;              // n triangle records, leaf order (output)
    uint32_t* perm;           // n: scene index of leaf-order triangle k (output)
    float4* nodes;            // room for max_nodes 4-wide nodes (output), depth-first order
    int32_t* level_nodes;     // room for max_nodes: node indices level by level, deepest level first (output)
    size_t max_nodes;         // at least max(n - 1, 1)
    void* scratch;            // gpu_build_bytes(n) bytes
    size_t scratch_bytes;
    uint32_t* host_words;     // 2 words of pinned host memory
    int n;
};
struct GpuBuildResult {
    int nodes = 0;
    int stack_need = 0;
    std::vector<int> level_count;  // 4-wide nodes per level, root level first
};
size_t gpu_build_bytes(int n);
// Returns a hipError_t code; waits on the stream once per clustering round
// and once per output level.
int gpu_build(const GpuBuildArgs& a, GpuBuildResult& r, hipStream_t st);
// dst[k] = src[perm[k]] for n triangle records.
void launch_gather_tris(const float4* src, const uint32_t* perm, int n, float4* dst, hipStream_t st);

// Bottom-up refit of the nodes listed in `order` (one tree level, deepest first).
void launch_refit_level(float4* nodes, const float4* geo, const int32_t* level_nodes, int count, hipStream_t st);

}  // namespace rmh
