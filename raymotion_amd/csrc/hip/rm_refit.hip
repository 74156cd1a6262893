// Bottom-up refit of the 4-wide tree after vertices moved (the GPU side of the
// reference's refit-or-rebuild policy, metal.mm:194-231). One launch per tree
// level, deepest level first; each lane recomputes the four child boxes of one
// node from its leaves' vertices or from the (already refit) child nodes.
#include "rm_kernels.h"

namespace rmh {
namespace {

__global__ __launch_bounds__(256) void k_refit(float4* __restrict__ nodes, const float4* __restrict__ geo,
                                                const int32_t* __restrict__ level_nodes, int count) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    float4* n = nodes + size_t(level_nodes[i]) * kNodeWords;
    int ref[kWide], kind[kWide];
    read_node_links(n, ref, kind);
    float lo[3][kWide], hi[3][kWide];
    read_node_boxes(n, lo, hi);
    for (int c = 0; c < kWide; ++c) {
        if (kind[c] < 0) continue;
        float bl[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
        float bh[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
        if (kind[c] == 0) {
            // An inner child: the union of its (already refit) child boxes.
            const float4* m = nodes + size_t(ref[c]) * kNodeWords;
            int mref[kWide], mkind[kWide];
            read_node_links(m, mref, mkind);
            float ml[3][kWide], mh[3][kWide];
            read_node_boxes(m, ml, mh);
            for (int k = 0; k < kWide; ++k)
                if (mkind[k] >= 0)
                    for (int a = 0; a < 3; ++a) {
                        bl[a] = fminf(bl[a], ml[a][k]);
                        bh[a] = fmaxf(bh[a], mh[a][k]);
                    }
        } else {
            for (int t = 0; t < kind[c]; ++t) {
                const float4* g = geo + size_t(ref[c] + t) * kTriGeoWords;
                const float4 g0 = g[0], g1 = g[1], g2 = g[2];
                const float vx[3] = {g0.x, g0.w, g1.z}, vy[3] = {g0.y, g1.x, g1.w}, vz[3] = {g0.z, g1.y, g2.x};
                for (int k = 0; k < 3; ++k) {
                    bl[0] = fminf(bl[0], vx[k]); bh[0] = fmaxf(bh[0], vx[k]);
                    bl[1] = fminf(bl[1], vy[k]); bh[1] = fmaxf(bh[1], vy[k]);
                    bl[2] = fminf(bl[2], vz[k]); bh[2] = fmaxf(bh[2], vz[k]);
                }
            }
        }
        for (int a = 0; a < 3; ++a) {
            lo[a][c] = bl[a];
            hi[a][c] = bh[a];
        }
    }
    float4 w[kNodeWords];
    for (int k = 0; k < kNodeWords; ++k) w[k] = n[k];
    write_node(w, lo, hi, ref, kind);
    for (int k = 0; k < kNodeWords; ++k) n[k] = w[k];
}

}  // namespace

void launch_refit_level(float4* nodes, const float4* geo, const int32_t* level_nodes, int count, hipStream_t st) {
    if (count <= 0) return;
    hipLaunchKernelGGL(k_refit, dim3((count + 255) / 256), dim3(256), 0, st, nodes, geo, level_nodes, count);
}

}  // namespace rmh
