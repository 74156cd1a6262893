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
    const float4 rf = n[6], kd = n[7];
    const int ref[4] = {__float_as_int(rf.x), __float_as_int(rf.y), __float_as_int(rf.z), __float_as_int(rf.w)};
    const int kind[4] = {__float_as_int(kd.x), __float_as_int(kd.y), __float_as_int(kd.z), __float_as_int(kd.w)};
    float lo[3][4], hi[3][4];
    {
        const float4 w0 = n[0], w1 = n[1], w2 = n[2], w3 = n[3], w4 = n[4], w5 = n[5];
        const float4 w[6] = {w0, w1, w2, w3, w4, w5};
        for (int a = 0; a < 3; ++a) {
            lo[a][0] = w[2 * a].x; lo[a][1] = w[2 * a].y; lo[a][2] = w[2 * a].z; lo[a][3] = w[2 * a].w;
            hi[a][0] = w[2 * a + 1].x; hi[a][1] = w[2 * a + 1].y; hi[a][2] = w[2 * a + 1].z; hi[a][3] = w[2 * a + 1].w;
        }
    }
    for (int c = 0; c < 4; ++c) {
        if (kind[c] < 0) continue;
        float bl[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
        float bh[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
        if (kind[c] == 0) {
            const float4* m = nodes + size_t(ref[c]) * kNodeWords;
            const float4 mk = m[7];
            const int mkind[4] = {__float_as_int(mk.x), __float_as_int(mk.y), __float_as_int(mk.z), __float_as_int(mk.w)};
            for (int a = 0; a < 3; ++a) {
                const float4 l = m[2 * a], h = m[2 * a + 1];
                const float lv[4] = {l.x, l.y, l.z, l.w}, hv[4] = {h.x, h.y, h.z, h.w};
                for (int k = 0; k < 4; ++k)
                    if (mkind[k] >= 0) {
                        bl[a] = fminf(bl[a], lv[k]);
                        bh[a] = fmaxf(bh[a], hv[k]);
                    }
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
    for (int a = 0; a < 3; ++a) {
        n[2 * a] = make_float4(lo[a][0], lo[a][1], lo[a][2], lo[a][3]);
        n[2 * a + 1] = make_float4(hi[a][0], hi[a][1], hi[a][2], hi[a][3]);
    }
}

}  // namespace

void launch_refit_level(float4* nodes, const float4* geo, const int32_t* level_nodes, int count, hipStream_t st) {
    if (count <= 0) return;
    hipLaunchKernelGGL(k_refit, dim3((count + 255) / 256), dim3(256), 0, st, nodes, geo, level_nodes, count);
}

}  // namespace rmh
