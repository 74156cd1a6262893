// Path-replay adjoint of the path-tracing megakernel (render_paths,
// pathtrace.metal:739-815 with pathtrace at 552-728): gradients of an image
// loss with respect to per-material albedo, roughness and emission scale.
//
// Sampling is detached: pdfs, MIS weights and roulette probabilities are held
// constant, and only the integrand's BSDF values and emission are
// differentiated. Every sample is replayed from its saved RNG state (no path
// vertices are stored; per sample stream the round in flight keeps 32 bytes):
// a primal pass gives the sample's radiance C and, in one 64-bit word, the
// outcomes of its light samples' shadow tests; a second pass re-traces the
// same path without re-tracing those shadow rays (on scenes whose surfaces
// are all opaque a shadow ray draws no random numbers, so skipping it keeps
// the replay's draws; elsewhere it is traced again), and at each vertex
//   emission      dC/d scale   = thr * base_emission * w
//   NEE           dC/d albedo  = thr * Le * k * df/d albedo      (at the light)
//                 dC/d rough   = thr * Le * k * df/d alpha2 * 4 r^3
//                 dC/d scale_l = thr * base_l * f * k
//   scattering    dC/d p       = R * (df/dp) / f   with R = C - (radiance so far),
//                                                   the part carried by this bounce
//   rough glass   dC/d rough   = R . 1 * d log w / d alpha2 * 4 r^3
// each dotted with dL/dC = dL/d(pixel) / n (zero for a channel the per-sample
// clamp cut off).
#include "rm_device.h"
#include "rm_kernels.h"

#ifndef RM_ADJ_BLOCKS
#define RM_ADJ_BLOCKS 3
#endif

namespace rmh {
namespace {

// Bounce-loop hooks that accumulate gradients into the block's LDS table.
struct AdjointHooks {
    const DevScene& sc;
    const AdjointArgs& A;
    double* g;     // LDS: 5 per group
    V3 C, dC;      // primal radiance of the sample and dL/dC
    uint64_t vis;  // the primal's shadow-test outcomes, one bit per light sample
    bool reuse_vis;
    int nvis = 0;

    template <class Tr, class NS>
    __device__ __forceinline__ bool visible(Tr& t, const NS& ns) {
        const int i = nvis++;
        if (reuse_vis && i < 64) return ((vis >> i) & 1u) != 0;
        return ns.shadow && !t.occluded(ns.o, ns.d, ns.dist);
    }

    __device__ void add(int grp, int k, float v) {
        if (v != 0.0f) atomicAdd(&g[grp * 5 + k], double(v));
    }
    // d alpha2 / d roughness for the material's raw roughness r
    // (alpha2 = max(min(max(r, 0), 1), 0.02)^4).
    static __device__ float drough(float r) { return (r > 0.02f && r < 1.0f) ? 4.0f * r * r * r : 0.0f; }
    // d(hit albedo) / d(material albedo): the texel factor, 0 for the azimuth shader.
    __device__ V3 albedo_factor(const Surface& hit) const {
        if (hit.shader == 1) return {0.0f, 0.0f, 0.0f};
        const float4 m = sc.mats[size_t(sc.matid[hit.tri]) * kMatWords];
        return {m.x != 0.0f ? hit.ar / m.x : 0.0f, m.y != 0.0f ? hit.ag / m.y : 0.0f,
                m.z != 0.0f ? hit.ab / m.z : 0.0f};
    }

    __device__ void emission(const Surface& hit, float tr, float tg, float tb, float w) {
        const int grp = A.tri_group[hit.tri];
        if (grp < 0) return;
        const float* b = A.group_base + 3 * grp;
        add(grp, 4, (dC.x * tr * b[0] + dC.y * tg * b[1] + dC.z * tb * b[2]) * w);
    }

    template <class NS>
    __device__ void nee(const Surface& hit, const NS& ns, float tr, float tg, float tb, float wox, float woy, float woz,
                        float nx, float ny, float nz, float alpha2, float metallic) {
        if (ns.light < 0) return;
        const int gl = A.light_group[ns.light];
        if (gl >= 0) {
            const float* b = A.group_base + 3 * gl;
            add(gl, 4, (dC.x * tr * b[0] * ns.er + dC.y * tg * b[1] * ns.eg + dC.z * tb * b[2] * ns.eb) * ns.k);
        }
        const int grp = A.tri_group[hit.tri];
        if (grp < 0) return;
        const BsdfGrad d = Tracer<32>::eval_bsdf_grad(ns.d.x, ns.d.y, ns.d.z, wox, woy, woz, nx, ny, nz, alpha2,
                                                      metallic, hit.ar, hit.ag, hit.ab);
        const float wr = dC.x * tr * ns.ler * ns.k, wg = dC.y * tg * ns.leg * ns.k, wb = dC.z * tb * ns.leb * ns.k;
        const V3 af = albedo_factor(hit);
        add(grp, 0, wr * d.da_r * af.x);
        add(grp, 1, wg * d.da_g * af.y);
        add(grp, 2, wb * d.da_b * af.z);
        add(grp, 3, (wr * d.dl_r + wg * d.dl_g + wb * d.dl_b) * drough(hit.rough));
    }

    __device__ void scatter(const Surface& hit, const Bsdf& e, float ix, float iy, float iz, float wox, float woy,
                            float woz, float nx, float ny, float nz, float alpha2, float metallic, float ar_, float ag_,
                            float ab_) {
        const int grp = A.tri_group[hit.tri];
        if (grp < 0) return;
        const float rr = dC.x * (C.x - ar_), rg = dC.y * (C.y - ag_), rb = dC.z * (C.z - ab_);
        if (rr == 0.0f && rg == 0.0f && rb == 0.0f) return;
        const BsdfGrad d =
            Tracer<32>::eval_bsdf_grad(ix, iy, iz, wox, woy, woz, nx, ny, nz, alpha2, metallic, hit.ar, hit.ag, hit.ab);
        const float qr = e.r != 0.0f ? rr / e.r : 0.0f, qg = e.g != 0.0f ? rg / e.g : 0.0f,
                    qb = e.b != 0.0f ? rb / e.b : 0.0f;
        const V3 af = albedo_factor(hit);
        add(grp, 0, qr * d.da_r * af.x);
        add(grp, 1, qg * d.da_g * af.y);
        add(grp, 2, qb * d.da_b * af.z);
        add(grp, 3, (qr * d.dl_r + qg * d.dl_g + qb * d.dl_b) * drough(hit.rough));
    }

    __device__ void dielectric(const Surface& hit, const DielectricSample& ds, float ar_, float ag_, float ab_) {
        if (ds.dlw_da2 == 0.0f) return;
        const int grp = A.tri_group[hit.tri];
        if (grp < 0) return;
        const float rsum = dC.x * (C.x - ar_) + dC.y * (C.y - ag_) + dC.z * (C.z - ab_);
        add(grp, 3, rsum * ds.dlw_da2 * drough(hit.rough));
    }
};

// The sample streams of render_paths_simd (pathtrace.metal:773-815): lane l
// of a pixel draws from pixel_seed ^ l * 0x94D049BB133111EB and takes the
// pixel's samples l, l + lanes, ... One round r runs sample r * lanes + l of
// every stream as two launches: the primal kernel traces it and stores its
// radiance, start and end RNG states and shadow-test outcomes; the replay
// kernel re-traces it from the start state with the gradient hooks. Each
// kernel holds one path loop, so neither carries the other's registers.
__device__ __forceinline__ uint64_t stream_seed(const FrameParams& P, uint32_t x, uint32_t y, uint32_t lane,
                                                int lanes) {
    uint64_t s = pixel_seed(P.seed, x, y);
    if (lanes > 1) s ^= uint64_t(lane) * 0x94D049BB133111EBULL;
    return s;
}

// Resident blocks per CU both kernels are register-limited to: 3 (3 waves
// per SIMD, at most 168 VGPRs), as the forward megakernel runs; the LDS
// stacks of 3 blocks fit up to S = 48.
constexpr int kAdjointBlocksPerCU = RM_ADJ_BLOCKS;

template <int S>
__global__ __launch_bounds__(kBlock, kAdjointBlocksPerCU) void k_adjoint_primal(DevScene sc, FrameParams P, AdjointArgs A, int lanes,
                                                            int round, unsigned long long* cnt) {
    __shared__ uint32_t stack[lds_stack_entries<S>() * kBlock];
    const uint32_t gid = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t pixel = gid / uint32_t(lanes), lane = gid % uint32_t(lanes);
    if (pixel >= uint32_t(P.width) * uint32_t(P.height)) return;
    if (round * lanes + int(lane) >= P.spp) return;
    const uint32_t x = pixel % uint32_t(P.width), y = pixel / uint32_t(P.width);
    Tracer<S> pt(sc, P, stack + threadIdx.x);
    if (round == 0) pt.rng.seed(stream_seed(P, x, y, lane, lanes));
    else pt.rng.state = A.record[gid].y;
    const uint32_t start = pt.rng.state;
    float sx, sy;
    pt.jitter(x, y, sx, sy);
    VisibilityRecorder rec;
    const V3 c = pt.pathtrace_t({P.cam.x, P.cam.y, P.cam.z}, pt.primary_dir(sx, sy, false), P.bounces, rec);
    A.primal[gid] = make_float4(c.x, c.y, c.z, 0.0f);
    A.record[gid] = make_uint4(start, pt.rng.state, uint32_t(rec.mask), uint32_t(rec.mask >> 32));
    pt.flush(cnt, kCntProbe);
}

template <int S>
__global__ __launch_bounds__(kBlock, kAdjointBlocksPerCU) void k_adjoint_replay(DevScene sc, FrameParams P, AdjointArgs A, int lanes,
                                                            int round, unsigned long long* cnt) {
    __shared__ uint32_t stack[lds_stack_entries<S>() * kBlock];
    // The block's gradient table, 5 per group, sized at launch to the scene's
    // groups (640 B at 16 groups instead of 5 KiB for the largest count).
    extern __shared__ double grad[];
    for (int i = threadIdx.x; i < A.ngroup * 5; i += kBlock) grad[i] = 0.0;
    __syncthreads();
    const uint32_t gid = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t pixel = gid / uint32_t(lanes), lane = gid % uint32_t(lanes);
    if (pixel < uint32_t(P.width) * uint32_t(P.height) && round * lanes + int(lane) < P.spp) {
        const uint32_t x = pixel % uint32_t(P.width), y = pixel / uint32_t(P.width);
        Tracer<S> pt(sc, P, stack + threadIdx.x);
        const uint4 r = A.record[gid];
        const float4 c4 = A.primal[gid];
        const V3 c{c4.x, c4.y, c4.z};
        pt.rng.state = r.x;
        const float fn = float(P.spp);
        V3 d{A.d_img[3 * pixel] / fn, A.d_img[3 * pixel + 1] / fn, A.d_img[3 * pixel + 2] / fn};
        if (P.clamp_value > 0.0f) {  // min(c, clamp) passes no gradient where it clamps
            if (c.x > P.clamp_value) d.x = 0.0f;
            if (c.y > P.clamp_value) d.y = 0.0f;
            if (c.z > P.clamp_value) d.z = 0.0f;
        }
        float sx, sy;
        pt.jitter(x, y, sx, sy);
        AdjointHooks hk{sc, A, grad, c, d, uint64_t(r.z) | (uint64_t(r.w) << 32), sc.opaque != 0};
        const V3 c2 = pt.pathtrace_t({P.cam.x, P.cam.y, P.cam.z}, pt.primary_dir(sx, sy, false), P.bounces, hk);
        if (__float_as_uint(c2.x) != __float_as_uint(c.x) || __float_as_uint(c2.y) != __float_as_uint(c.y) ||
            __float_as_uint(c2.z) != __float_as_uint(c.z) || pt.rng.state != r.y)
            atomicAdd(A.mismatch, 1ull);
        pt.flush(cnt, kCntProbe);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < A.ngroup * 5; i += kBlock)
        if (grad[i] != 0.0) atomicAdd(&A.grad[i], grad[i]);
}

}  // namespace

void launch_adjoint(int stack, int lanes, const DevScene& sc, const FrameParams& P, const AdjointArgs& A,
                    unsigned long long* cnt, hipStream_t st) {
    const size_t threads = size_t(P.width) * size_t(P.height) * size_t(lanes);
    const dim3 grid(unsigned((threads + kBlock - 1) / kBlock)), block(kBlock);
    const int rounds = (P.spp + lanes - 1) / lanes;
    with_stack(stack, [&](auto cap) {
        for (int r = 0; r < rounds; ++r) {
            hipLaunchKernelGGL((k_adjoint_primal<decltype(cap)::value>), grid, block, 0, st, sc, P, A, lanes, r, cnt);
            hipLaunchKernelGGL((k_adjoint_replay<decltype(cap)::value>), grid, block,
                               size_t(A.ngroup) * 5 * sizeof(double), st, sc, P, A, lanes, r, cnt);
        }
    });
}

}  // namespace rmh
