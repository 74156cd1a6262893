// gfx950 kernels of the rendering backend, one per pass of the reference's
// frame (src/engine/pathtrace.metal:739-1363), plus the traversal probe.
// All passes are 1-D over pixels with 256-thread blocks (4 wave64s); every
// tracing lane owns a column of an LDS stack array of S entries.
#include "rm_device.h"
#include "rm_kernels.h"

namespace rmh {

namespace {

__device__ __forceinline__ V3 clamp3(V3 c, float cl) {
    return cl > 0.0f ? vmin(c, V3{cl, cl, cl}) : c;
}

// render_paths_simd<LANES> (pathtrace.metal:773-815): LANES lanes per pixel,
// each with its own stream, combined by an xor butterfly per block of samples.
// (One lane per pixel runs in the persistent engine, rm_wave.hip.)
template <int S, int LANES>
__global__ __launch_bounds__(kBlock) void k_render_paths(DevScene sc, FrameParams P, uint8_t* __restrict__ out,
                                                          float4* __restrict__ hdr, unsigned long long* cnt) {
    __shared__ uint32_t stack[lds_stack_entries<S>() * kBlock];
    const uint32_t gid = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t pixel = gid / LANES, lane = gid % LANES;
    const uint32_t npx = uint32_t(P.width) * uint32_t(P.height);
    if (pixel >= npx) return;  // lane groups are aligned, so a group leaves together
    const uint32_t x = pixel % uint32_t(P.width), y = pixel / uint32_t(P.width);
    Tracer<S> pt(sc, P, stack + threadIdx.x);
    uint64_t s = pixel_seed(P.seed, x, y);
    s ^= uint64_t(lane) * 0x94D049BB133111EBULL;
    pt.rng.seed(s);
    float sr = 0.0f, sg = 0.0f, sb = 0.0f, sum_l = 0.0f, sum_l2 = 0.0f;
    int n = 0;
    for (int base = 0; base < P.spp; base += LANES) {
        V3 c{0.0f, 0.0f, 0.0f};
        if (base + int(lane) < P.spp) {
            float sx, sy;
            pt.jitter(x, y, sx, sy);
            c = clamp3(pt.sample_sensor(sx, sy), P.clamp_value);
        }
        float lum = c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f;
        float lum2 = lum * lum;
#pragma unroll
        for (int mask = 1; mask < LANES; mask *= 2) {
            c.x += __shfl_xor(c.x, mask, LANES);
            c.y += __shfl_xor(c.y, mask, LANES);
            c.z += __shfl_xor(c.z, mask, LANES);
            lum += __shfl_xor(lum, mask, LANES);
            lum2 += __shfl_xor(lum2, mask, LANES);
        }
        sr += c.x; sg += c.y; sb += c.z;
        sum_l += lum;
        sum_l2 += lum2;
        n = min(base + LANES, P.spp);
        if (P.adapt_min > 0 && n >= P.adapt_min && n > 1 && n % P.adapt_step == 0) {
            const float mean = sum_l / float(n);
            const float var = fmaxf(0.0f, (sum_l2 - sum_l * mean) / float(n - 1));
            const float th = P.adapt_rel * mean + P.adapt_abs;
            if (var < th * th) break;
        }
    }
    pt.flush(cnt, kCntMega);
    if (lane == 0) {
        const float mr = sr / float(n), mg = sg / float(n), mb = sb / float(n);
        out[3 * pixel] = tonemap8(mr);
        out[3 * pixel + 1] = tonemap8(mg);
        out[3 * pixel + 2] = tonemap8(mb);
        if (hdr) hdr[pixel] = make_float4(mr, mg, mb, 0.0f);
    }
}

__device__ __forceinline__ bool reproject(const FrameParams& P, V3 p, uint32_t& hidx) {  // pathtrace.metal:957-966
    const rm_camera& c = P.prev;
    const float pdx = p.x - c.x, pdy = p.y - c.y, pdz = p.z - c.z;
    const float ccx = c.m0 * pdx + c.m3 * pdy + c.m6 * pdz;
    const float ccy = c.m1 * pdx + c.m4 * pdy + c.m7 * pdz;
    const float ccz = c.m2 * pdx + c.m5 * pdy + c.m8 * pdz;
    if (!(ccz > 1e-6f)) return false;
    const float ppx = c.focal * ccx / ccz + float(P.width) * 0.5f;
    const float ppy = float(P.height) * 0.5f - c.focal * ccy / ccz;
    const int hx = int(floorf(ppx)), hy = int(floorf(ppy));
    if (!(hx >= 0 && hy >= 0 && hx < P.width && hy < P.height)) return false;
    hidx = uint32_t(hy) * uint32_t(P.width) + uint32_t(hx);
    return true;
}

// restir_spatial (pathtrace.metal:919-1054): temporal then spatial reuse.
template <int S>
__global__ __launch_bounds__(kBlock) void k_spatial(DevScene sc, FrameParams P, GBufferSoA gb, ReservoirSoA ri,
                                                     ReservoirSoA ro_, GBufferSoA hg, ReservoirSoA hr,
                                                     unsigned long long* cnt) {
    __shared__ uint32_t stack[lds_stack_entries<S>() * kBlock];
    const uint32_t id = blockIdx.x * kBlock + threadIdx.x;
    if (id >= uint32_t(P.width) * uint32_t(P.height)) return;
    const uint32_t x = id % uint32_t(P.width), y = id / uint32_t(P.width);
    const float4 i0 = ri.r0[id], i1 = ri.r1[id], i2 = ri.r2[id], i3 = ri.r3[id];
    const float4 g0 = gb.p0[id];
    int Sn = P.restir_spatial;
    if (Sn < 0) Sn = 0;
    if (Sn > 8) Sn = 8;
    // Pixels that merge nothing keep their initial reservoir; each output is
    // written once (an early copy overwritten after the shadow rays would
    // reach HBM twice).
    if (ibits(g0.w) != 1 || __float_as_uint(i3.z) == 0u || Sn == 0) {
        ro_.r0[id] = i0; ro_.r1[id] = i1; ro_.r2[id] = i2; ro_.r3[id] = i3;
        return;
    }
    float radius = P.restir_radius;
    if (!(radius > 0.0f)) radius = 16.0f;
    int mcap = P.restir_mcap;
    if (mcap < 1) mcap = 1;
    Tracer<S> pt(sc, P, stack + threadIdx.x);
    pt.rng.seed(P.seed ^ (uint64_t(id + 1) * 0x9E3779B97F4A7C15ULL) ^ kSpatialSalt);
    const float4 g1 = gb.p1[id], g2 = gb.p2[id], g3 = gb.p3[id];
    const V3 cpos = xyz(g0), cnor = xyz(g1), cwo = xyz(g3), calb = xyz(g2);
    float rough = g1.w;
    if (rough < 0.0f) rough = 0.0f;
    if (rough > 1.0f) rough = 1.0f;
    float metallic = g2.w;
    if (metallic < 0.0f) metallic = 0.0f;
    if (metallic > 1.0f) metallic = 1.0f;
    float r2 = rough;
    if (r2 < 0.02f) r2 = 0.02f;
    const float alpha2 = (r2 * r2) * (r2 * r2);
    float spec_prob = 0.05f + 0.95f * metallic;
    if (spec_prob > 0.98f) spec_prob = 0.98f;
    if (metallic >= 0.999f) spec_prob = 1.0f;
    float cur_sumWs = i2.w, cur_sumWu = i3.x, cur_ph = i0.w;
    uint32_t cur_M = __float_as_uint(i3.y);
    float4 cur0 = i0, cur1 = i1, cur2 = i2;  // light pos / normal+alpha / emission
    const V3 so{cpos.x + cnor.x * 0.002f, cpos.y + cnor.y * 0.002f, cpos.z + cnor.z * 0.002f};
    uint32_t hidx;
    if (P.has_history != 0 && reproject(P, cpos, hidx) && ibits(hg.p0[hidx].w) == 1 &&
        __float_as_uint(hr.r3[hidx].z) != 0u) {
        const float4 hp0 = hg.p0[hidx], hp1 = hg.p1[hidx];
        const V3 hp = xyz(hp0), hn = xyz(hp1);
        const float ddx = cpos.x - hp.x, ddy = cpos.y - hp.y, ddz = cpos.z - hp.z;
        const float dd = sqrtf(ddx * ddx + ddy * ddy + ddz * ddz);
        const float dcx = cpos.x - P.cam.x, dcy = cpos.y - P.cam.y, dcz = cpos.z - P.cam.z;
        const float dc = sqrtf(dcx * dcx + dcy * dcy + dcz * dcz);
        const float4 h0 = hr.r0[hidx], h1 = hr.r1[hidx], h2 = hr.r2[hidx], h3 = hr.r3[hidx];
        const float hsumWs = h2.w, hsumWu = h3.x, hph = h0.w;
        const uint32_t hM = __float_as_uint(h3.y);
        if (dd < 0.02f * dc + 1e-3f && (cnor.x * hn.x + (cnor.y * hn.y + cnor.z * hn.z)) > 0.9f && hsumWu > 0.0f &&
            hM > 0u && hph > 0.0f) {
            const V3 hlp = xyz(h0), hln = xyz(h1), hle = xyz(h2);
            const float halpha = h1.w;
            const float ldx = hlp.x - cpos.x, ldy = hlp.y - cpos.y, ldz = hlp.z - cpos.z;
            const float dist = sqrtf(ldx * ldx + ldy * ldy + ldz * ldz);
            if (dist > 0.003f) {
                const float inv = 1.0f / dist;
                const float dxn = ldx * inv, dyn = ldy * inv, dzn = ldz * inv;
                const float lcos = -(hln.x * dxn + (hln.y * dyn + hln.z * dzn));
                const float scos = cnor.x * dxn + (cnor.y * dyn + cnor.z * dzn);
                if (lcos > 1e-8f && scos > 1e-8f) {
                    const Bsdf e = Tracer<S>::eval_bsdf(dxn, dyn, dzn, cwo.x, cwo.y, cwo.z, cnor.x, cnor.y, cnor.z,
                                                        alpha2, metallic, spec_prob, calb.x, calb.y, calb.z);
                    const float ph = lum3(hle.x * e.r * scos * halpha, hle.y * e.g * scos * halpha,
                                          hle.z * e.b * scos * halpha);
                    if (ph > 0.0f) {
                        const float wu = ph * hsumWu / hph;
                        const bool occ = pt.occluded(so, {dxn, dyn, dzn}, dist - 0.003f);
                        const float ws = occ ? 0.0f : ph * hsumWs / hph;
                        const float sum_s = cur_sumWs + ws, sum_u = cur_sumWu + wu;
                        uint32_t newM = cur_M + hM;
                        if (newM > uint32_t(mcap)) newM = uint32_t(mcap);
                        const float r = pt.urand();
                        if (sum_s > 0.0f && r * sum_s < ws) {
                            cur0 = h0; cur1 = h1; cur2 = h2;
                            cur_ph = ph;
                        }
                        cur_sumWs = sum_s; cur_sumWu = sum_u; cur_M = newM;
                    }
                }
            }
        }
    }
    for (int s = 0; s < Sn; ++s) {
        const float rx = pt.urand() * 2.0f - 1.0f, ry = pt.urand() * 2.0f - 1.0f;
        const int nx = int(float(x) + rx * radius + 0.5f), ny = int(float(y) + ry * radius + 0.5f);
        if (nx < 0 || ny < 0 || nx >= P.width || ny >= P.height) continue;
        if (nx == int(x) && ny == int(y)) continue;
        const uint32_t nidx = uint32_t(ny) * uint32_t(P.width) + uint32_t(nx);
        const float4 n3 = ri.r3[nidx];
        if (ibits(gb.p0[nidx].w) != 1 || __float_as_uint(n3.z) == 0u) continue;
        const V3 nnor = xyz(gb.p1[nidx]);
        if (cnor.x * nnor.x + (cnor.y * nnor.y + cnor.z * nnor.z) < 0.9f) continue;
        const float4 n0 = ri.r0[nidx], n1 = ri.r1[nidx], n2 = ri.r2[nidx];
        const float nbr_sumWs = n2.w, nbr_sumWu = n3.x, nbr_ph = n0.w;
        const uint32_t nbr_M = __float_as_uint(n3.y);
        if (!(nbr_sumWu > 0.0f) || nbr_M == 0u || !(nbr_ph > 0.0f)) continue;
        const V3 nlp = xyz(n0), nln = xyz(n1), nle = xyz(n2);
        const float nalpha = n1.w;
        const float ldx = nlp.x - cpos.x, ldy = nlp.y - cpos.y, ldz = nlp.z - cpos.z;
        const float dist = sqrtf(ldx * ldx + ldy * ldy + ldz * ldz);
        if (!(dist > 0.003f)) continue;
        const float inv = 1.0f / dist;
        const float dxn = ldx * inv, dyn = ldy * inv, dzn = ldz * inv;
        const float lcos = -(nln.x * dxn + (nln.y * dyn + nln.z * dzn));
        const float scos = cnor.x * dxn + (cnor.y * dyn + cnor.z * dzn);
        if (!(lcos > 1e-8f && scos > 1e-8f)) continue;
        const Bsdf e = Tracer<S>::eval_bsdf(dxn, dyn, dzn, cwo.x, cwo.y, cwo.z, cnor.x, cnor.y, cnor.z, alpha2,
                                            metallic, spec_prob, calb.x, calb.y, calb.z);
        const float ph_cur = lum3(nle.x * e.r * scos * nalpha, nle.y * e.g * scos * nalpha, nle.z * e.b * scos * nalpha);
        if (!(ph_cur > 0.0f)) continue;
        const float w_nbr_u = ph_cur * nbr_sumWu / nbr_ph;
        if (!(w_nbr_u > 0.0f)) continue;
        const bool occ = pt.occluded(so, {dxn, dyn, dzn}, dist - 0.003f);
        const float w_nbr_s = occ ? 0.0f : ph_cur * nbr_sumWs / nbr_ph;
        const float sum_s = cur_sumWs + w_nbr_s, sum_u = cur_sumWu + w_nbr_u;
        if (!(sum_u > 0.0f) || !(sum_s >= 0.0f)) continue;
        const float r = pt.urand();
        uint32_t newM = cur_M + nbr_M;
        if (newM > uint32_t(mcap)) newM = uint32_t(mcap);
        if (sum_s > 0.0f && r * sum_s < w_nbr_s) {
            cur0 = n0; cur1 = n1; cur2 = n2;
            cur_ph = ph_cur;
        }
        cur_sumWs = sum_s; cur_sumWu = sum_u; cur_M = newM;
    }
    pt.flush(cnt, kCntSpatial);
    ro_.r0[id] = make_float4(cur0.x, cur0.y, cur0.z, cur_ph);
    ro_.r1[id] = cur1;
    ro_.r2[id] = make_float4(cur2.x, cur2.y, cur2.z, cur_sumWs);
    ro_.r3[id] = make_float4(cur_sumWu, __uint_as_float(cur_M), __uint_as_float(1u), 0.0f);
}

__device__ __forceinline__ float luma(float4 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

// denoise_temporal (pathtrace.metal:1263-1314)
__global__ __launch_bounds__(kBlock) void k_temporal(FrameParams P, GBufferSoA gb, GBufferSoA hg,
                                                      const float4* __restrict__ in, const float4* __restrict__ hist,
                                                      float4* __restrict__ out) {
    const uint32_t id = blockIdx.x * kBlock + threadIdx.x;
    if (id >= uint32_t(P.width) * uint32_t(P.height)) return;
    const uint32_t x = id % uint32_t(P.width), y = id / uint32_t(P.width);
    const float4 cur = in[id];
    const float4 g0 = gb.p0[id];
    float4 outc = cur;
    if (P.has_history != 0 && ibits(g0.w) == 1) {
        float4 mn = cur, mx = cur;
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int nx = int(x) + dx, ny = int(y) + dy;
                if (nx < 0 || ny < 0 || nx >= P.width || ny >= P.height) continue;
                const float4 v = in[uint32_t(ny) * uint32_t(P.width) + uint32_t(nx)];
                mn = make_float4(fminf(mn.x, v.x), fminf(mn.y, v.y), fminf(mn.z, v.z), 0.0f);
                mx = make_float4(fmaxf(mx.x, v.x), fmaxf(mx.y, v.y), fmaxf(mx.z, v.z), 0.0f);
            }
        const V3 cpos = xyz(g0), cnor = xyz(gb.p1[id]);
        uint32_t hidx;
        if (reproject(P, cpos, hidx) && ibits(hg.p0[hidx].w) == 1) {
            const V3 hp = xyz(hg.p0[hidx]), hn = xyz(hg.p1[hidx]);
            const float ddx = cpos.x - hp.x, ddy = cpos.y - hp.y, ddz = cpos.z - hp.z;
            const float dd = sqrtf(ddx * ddx + ddy * ddy + ddz * ddz);
            const float dcx = cpos.x - P.cam.x, dcy = cpos.y - P.cam.y, dcz = cpos.z - P.cam.z;
            const float dc = sqrtf(dcx * dcx + dcy * dcy + dcz * dcz);
            if (dd < 0.02f * dc + 1e-3f && (cnor.x * hn.x + (cnor.y * hn.y + cnor.z * hn.z)) > 0.9f) {
                float4 h = hist[hidx];
                h.x = fminf(fmaxf(h.x, mn.x), mx.x);
                h.y = fminf(fmaxf(h.y, mn.y), mx.y);
                h.z = fminf(fmaxf(h.z, mn.z), mx.z);
                outc = make_float4(h.x + (cur.x - h.x) * 0.2f, h.y + (cur.y - h.y) * 0.2f, h.z + (cur.z - h.z) * 0.2f,
                                   0.0f);
            }
        }
    }
    out[id] = outc;
}

// denoise_atrous (pathtrace.metal:1315-1354): 5x5 taps at stride 2^step with
// normal, depth, luminance and albedo edge-stopping weights, each exponential
// weight as the reference writes it (quotient, then exp), with rm_expf: the
// host denoiser gives the same bits.
__global__ __launch_bounds__(kBlock) void k_atrous(FrameParams P, GBufferSoA gb, const float4* __restrict__ in,
                                                    float4* __restrict__ out, int step) {
    const uint32_t id = blockIdx.x * kBlock + threadIdx.x;
    if (id >= uint32_t(P.width) * uint32_t(P.height)) return;
    const uint32_t x = id % uint32_t(P.width), y = id / uint32_t(P.width);
    const float4 g0 = gb.p0[id];
    if (ibits(g0.w) == 0) {
        out[id] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    const float4 c0 = in[id];
    const V3 n0 = xyz(gb.p1[id]), p0 = xyz(g0), a0 = xyz(gb.p2[id]);
    const float l0 = luma(c0);
    const float dcx = p0.x - P.cam.x, dcy = p0.y - P.cam.y, dcz = p0.z - P.cam.z;
    const float dc = sqrtf(dcx * dcx + dcy * dcy + dcz * dcz);
    const float sz = 0.1f * dc + 1e-4f, sl = 0.3f * fmaxf(l0, 0.5f) + 1e-3f;
    const int stride = 1 << step;
    float sr = 0.0f, sg = 0.0f, sb = 0.0f, wsum = 0.0f;
    for (int dy = -2; dy <= 2; ++dy) {
        const int ny = int(y) + dy * stride;
        if (ny < 0 || ny >= P.height) continue;
        for (int dx = -2; dx <= 2; ++dx) {
            const int nx = int(x) + dx * stride;
            if (nx < 0 || nx >= P.width) continue;
            const uint32_t n = uint32_t(ny) * uint32_t(P.width) + uint32_t(nx);
            const float4 q0 = gb.p0[n];
            if (ibits(q0.w) == 0) continue;
            const float4 c = in[n];
            const V3 nn = xyz(gb.p1[n]), aa = xyz(gb.p2[n]);
            const float nd = nn.x * n0.x + (nn.y * n0.y + nn.z * n0.z);
            float wn = fmaxf(nd, 0.0f);
            wn = wn * wn; wn = wn * wn; wn = wn * wn; wn = wn * wn;
            const float ex = q0.x - p0.x, ey = q0.y - p0.y, ez = q0.z - p0.z;
            const float wz = rm_expf(-sqrtf(ex * ex + ey * ey + ez * ez) / sz);
            const float wl = rm_expf(-fabsf(luma(c) - l0) / sl);
            const float ax = aa.x - a0.x, ay = aa.y - a0.y, az = aa.z - a0.z;
            const float wa = rm_expf(-sqrtf(ax * ax + ay * ay + az * az) * 4.0f);  // / 0.25, exactly
            const float w = wn * wz * wl * wa;
            sr += c.x * w; sg += c.y * w; sb += c.z * w;
            wsum += w;
        }
    }
    out[id] = wsum > 1e-12f ? make_float4(sr / wsum, sg / wsum, sb / wsum, 0.0f) : c0;
}

// denoise_tonemap (pathtrace.metal:1355-1363); also exports the linear image.
__global__ __launch_bounds__(kBlock) void k_tonemap(int npx, const float4* __restrict__ in, uint8_t* __restrict__ rgb,
                                                     float* __restrict__ hdr3) {
    const int id = blockIdx.x * kBlock + threadIdx.x;
    if (id >= npx) return;
    const float4 c = in[id];
    rgb[3 * id] = tonemap8(c.x);
    rgb[3 * id + 1] = tonemap8(c.y);
    rgb[3 * id + 2] = tonemap8(c.z);
    if (hdr3) {
        hdr3[3 * id] = c.x;
        hdr3[3 * id + 1] = c.y;
        hdr3[3 * id + 2] = c.z;
    }
}

// Traversal probe: one closest- or any-hit query per lane.
template <int S>
__global__ __launch_bounds__(kBlock) void k_trace(DevScene sc, FrameParams P, int n, const float* __restrict__ o,
                                                   const float* __restrict__ d, const float* __restrict__ tmax, int any,
                                                   int32_t* __restrict__ tri, float4* __restrict__ tuv,
                                                   unsigned long long* cnt) {
    __shared__ uint32_t stack[lds_stack_entries<S>() * kBlock];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    Tracer<S> pt(sc, P, stack + threadIdx.x);
    const V3 ro{o[3 * i], o[3 * i + 1], o[3 * i + 2]}, rd{d[3 * i], d[3 * i + 1], d[3 * i + 2]};
    const Hit h = pt.intersect(ro, rd, tmax ? tmax[i] : 1e30f, any != 0);
    tri[i] = h.tri;
    tuv[i] = make_float4(h.t, h.u, h.v, 0.0f);
    pt.flush(cnt, kCntProbe);
}

// Zero fill of a small word array (in place of a memset node, which a caller
// capturing the stream into a graph cannot use).
__global__ __launch_bounds__(kBlock) void k_zero_u64(unsigned long long* p, int n) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

inline unsigned blocks_for(size_t n) { return unsigned((n + kBlock - 1) / kBlock); }

}  // namespace

#define RM_STACK_DISPATCH(stack, launch) \
    do {                                 \
        if ((stack) < 32) {              \
            constexpr int S = 32;        \
            launch;                      \
        } else if ((stack) < 48) {       \
            constexpr int S = 48;        \
            launch;                      \
        } else if ((stack) < 64) {       \
            constexpr int S = 64;        \
            launch;                      \
        } else {                         \
            constexpr int S = 256;       \
            launch;                      \
        }                                \
    } while (0)
static_assert(rmh::kLdsStack == 64 && rmh::kSpillStack == 256, "RM_STACK_DISPATCH instances");

void launch_render_paths(int stack, int lanes, const DevScene& sc, const FrameParams& P, uint8_t* out, float4* hdr,
                         unsigned long long* cnt, hipStream_t st) {
    const size_t threads = size_t(P.width) * size_t(P.height) * size_t(lanes);
    const dim3 grid(blocks_for(threads)), block(kBlock);
    RM_STACK_DISPATCH(stack, {
        if (lanes == 32) hipLaunchKernelGGL((k_render_paths<S, 32>), grid, block, 0, st, sc, P, out, hdr, cnt);
        else hipLaunchKernelGGL((k_render_paths<S, 8>), grid, block, 0, st, sc, P, out, hdr, cnt);
    });
}

void launch_spatial(int stack, const DevScene& sc, const FrameParams& P, const GBufferSoA& gb, const ReservoirSoA& ri,
                    const ReservoirSoA& ro, const GBufferSoA& hg, const ReservoirSoA& hr, unsigned long long* cnt,
                    hipStream_t st) {
    const dim3 grid(blocks_for(size_t(P.width) * size_t(P.height))), block(kBlock);
    RM_STACK_DISPATCH(stack, hipLaunchKernelGGL((k_spatial<S>), grid, block, 0, st, sc, P, gb, ri, ro, hg, hr, cnt));
}

void launch_temporal(const FrameParams& P, const GBufferSoA& gb, const GBufferSoA& hg, const float4* in,
                     const float4* hist, float4* out, hipStream_t st) {
    const dim3 grid(blocks_for(size_t(P.width) * size_t(P.height))), block(kBlock);
    hipLaunchKernelGGL(k_temporal, grid, block, 0, st, P, gb, hg, in, hist, out);
}

void launch_atrous(const FrameParams& P, const GBufferSoA& gb, const float4* in, float4* out, int step,
                   hipStream_t st) {
    const dim3 grid(blocks_for(size_t(P.width) * size_t(P.height))), block(kBlock);
    hipLaunchKernelGGL(k_atrous, grid, block, 0, st, P, gb, in, out, step);
}

void launch_tonemap(int npx, const float4* in, uint8_t* rgb, float* hdr3, hipStream_t st) {
    hipLaunchKernelGGL(k_tonemap, dim3(blocks_for(size_t(npx))), dim3(kBlock), 0, st, npx, in, rgb, hdr3);
}

void launch_zero_u64(unsigned long long* p, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_u64, dim3(blocks_for(size_t(n))), dim3(kBlock), 0, st, p, n);
}

void launch_trace(int stack, const DevScene& sc, const FrameParams& P, int n, const float* o, const float* d,
                  const float* tmax, int any, int32_t* tri, float4* tuv, unsigned long long* cnt, hipStream_t st) {
    const dim3 grid(blocks_for(size_t(n))), block(kBlock);
    RM_STACK_DISPATCH(stack, hipLaunchKernelGGL((k_trace<S>), grid, block, 0, st, sc, P, n, o, d, tmax, any, tri, tuv,
                                                cnt));
}

}  // namespace rmh
