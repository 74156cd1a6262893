// Persistent engines for the long-running passes: render_paths with one lane
// per pixel (pathtrace.metal:739-769), restir_shade (pathtrace.metal:1056-1258)
// and restir_generate (pathtrace.metal:817-917).
//
// The reference runs each pixel's whole sample loop inside one thread, so the
// threads of a wavefront sit at different places of their paths and trace
// their rays one at a time in divergent code. Here every lane keeps its
// pixel's state in registers and the kernel is a loop whose body traces
// exactly one ray per active lane (extension, shadow, or the next step of a
// stochastic-coverage walk), then advances each lane's state machine to its
// next ray. All active lanes of a wave therefore run the traversal together,
// and a lane whose pixel is finished takes the next pixel from a global
// counter (one atomic per wave), so waves stay full until the image is done.
// The per-pixel arithmetic and random-number order are those of the reference
// functions (shared with the other kernels through Tracer), which keeps every
// pixel's result bit-identical to a straight evaluation.
#include "rm_device.h"
#include "rm_kernels.h"

#include <map>
#include <mutex>
#include <utility>

namespace rmh {
namespace {

// kNee: the lane stands at a shading point and waits for k_persistent to run
// its light sampling together with the other waiting lanes of its wave.
// kDual: the lane's shadow ray and its next extension ray are both pending (the
// continuation was sampled right after the light sample); kShadowEnd: only the
// shadow ray is pending and the path ends after it.
enum Phase : int { kIdle = 0, kExt = 1, kShadow = 2, kNee = 3, kDual = 4, kShadowEnd = 5 };

// Index of the r-th (from 0) set bit of m; m has more than r set bits.
__device__ __forceinline__ int nth_set_bit(unsigned long long m, int r) {
    for (int i = 0; i < r; ++i) m &= m - 1ull;
    return __ffsll((long long)m) - 1;
}

// Position for the calling lane (when keep) in a list grown through *count,
// one atomic per wave; undefined without keep.
__device__ __forceinline__ unsigned wave_append_pos(bool keep, unsigned* count) {
    const unsigned long long m = __ballot(keep);
    if (!m) return 0u;
    const int leader = __ffsll((long long)m) - 1;
    unsigned base = 0;
    if (int(__lane_id()) == leader) base = atomicAdd(count, unsigned(__popcll(m)));
    base = __shfl(base, leader);
    return base + __builtin_amdgcn_mbcnt_hi(unsigned(m >> 32), __builtin_amdgcn_mbcnt_lo(unsigned(m), 0u));
}
enum Step : int { kTrace = 0, kPathEnd = 1, kPixelDone = 2, kWait = 3 };

struct PathLane {
    int phase;
    int pixel, kind;  // kind 1: ReSTIR indirect samples, 3: camera paths
    int s, n;
    float sr, sg, sb;     // sum of sample radiance
    float sl, sl2;        // adaptive-sampling luminance sums (megakernel)
    float dr, dg, db;     // emission + ReSTIR direct term of the primary (kind 1)
    float m_sumWu;
    uint32_t m_M;
    // path
    float tr, tg, tb, ar, ag, ab;
    V3 po, d;             // path vertex origin and direction of the current segment
    float last_pdf, prev_sumW;
    uint32_t prev_M;
    float pbr, pbg, pbb, pbcos;
    int prev_delta, hit, max_hits;
    bool clamp_neg;
    // ray being traced (stochastic-coverage walks advance along it)
    V3 ro, rd, base;
    float tmax, traveled, seq_dist;
    // vertex waiting for its shadow ray
    V3 p, pn;
    float h_ar, h_ag, h_ab, h_rough, h_alpha2, h_met, h_spec;
    float nr, ng, nb, nsumW;
    uint32_t nM;
    // shadow ray kept aside in phase kDual / kShadowEnd (it leaves from
    // p + 0.002 pn); nr, ng, nb then hold the light's contribution already
    // multiplied by the vertex's throughput
    V3 sd{0.0f, 0.0f, 0.0f};
    float sdist = 0.0f;
};

// Buffers a persistent pass reads or writes.
struct PassIO {
    GBufferSoA gb;
    ReservoirSoA res;
    float4* hdr;
    uint8_t* rgb;
    float* var;  // restir_shade: per-pixel sample variance of luminance (may be null)
};

template <int S, bool SHADE>
struct PathEngine {
    using Lane = PathLane;
    static constexpr int kPass = SHADE ? kCntShade : kCntMega;
    // Resident 256-lane blocks per CU the register allocation targets: 2
    // (waves per SIMD) without spills serves trees that stay in the caches
    // (327 K triangles, 163 K nodes: shade 10.1 ms against 10.2-10.3 ms at 3
    // with 196 B of spills per lane); PathEngineDeep below serves large trees.
    // The stack takes S KiB of LDS per block.
    static constexpr int kMinBlocks = 2;
    static constexpr int kLogicBlocks = 2;  // wavefront logic kernel (no LDS)
    static constexpr bool kDeferNee = true;
    // Lanes stop waiting for light-sampling batches once the queue is
    // exhausted (k_persistent): the shade pass's tail is its long camera-path
    // chains. The megakernel keeps waiting (no difference measured there).
    static constexpr bool kTailNoWait = SHADE;
    // Traversal inside the persistent loop with refill (k_persistent): not
    // kept here (C1 shade 10.1-10.2 ms against 9.9-10.0 ms with the call).
    static constexpr bool kFused = false;
    // Set by k_persistent: shade_hit leaves the light sampling to the loop
    // (phase kNee, step result kWait) instead of running it at once.
    bool defer_nee = false;
    // Set by k_persistent for scenes whose shadow rays draw no random numbers
    // (every surface opaque): after the light sample the continuation is
    // sampled at once, so the shadow ray and the extension ray are both
    // pending (phase kDual) and an idle lane of the wave can trace the shadow
    // ray in the same traversal round as the extension. The RNG draws and the
    // order of the radiance sums stay those of the reference's loop
    // (direct_lighting's draws, then the BSDF sample, then roulette; the
    // light's term is added before the next vertex's emission).
    bool dual_shadow = false;
    static constexpr bool kDualShadow = true;
    Tracer<S>& pt;
    const FrameParams& P;
    const GBufferSoA& gb;
    const ReservoirSoA& rm;
    float4* hdr;
    uint8_t* rgb;
    float* lum_var;
    __device__ PathEngine(Tracer<S>& t, const FrameParams& p, const PassIO& io)
        : pt(t), P(p), gb(io.gb), rm(io.res), hdr(io.hdr), rgb(io.rgb), lum_var(io.var) {}

    __device__ void write_var(int px, float v) {
        if (SHADE && lum_var) lum_var[px] = v;
    }
    __device__ void add_lum(Lane& L, V3 c) {  // luminance moments for lum_var
        const float lum = c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f;
        L.sl += lum;
        L.sl2 += lum * lum;
    }

    __device__ void set_ray(Lane& L, int phase, V3 o, V3 dir, float dist) {
        L.phase = phase;
        L.ro = o;
        L.base = o;
        L.rd = dir;
        L.tmax = dist;
        L.traveled = 0.0f;
        L.seq_dist = dist;
    }

    // Camera path (pathtrace from sample_sensor): megakernel and kind 3.
    __device__ int begin_camera_sample(Lane& L) {
        const uint32_t x = uint32_t(L.pixel) % uint32_t(P.width), y = uint32_t(L.pixel) / uint32_t(P.width);
        float sx, sy;
        pt.jitter(x, y, sx, sy);
        L.d = pt.primary_dir(sx, sy, false);
        L.po = V3{P.cam.x, P.cam.y, P.cam.z};
        L.tr = L.tg = L.tb = 1.0f;
        L.ar = L.ag = L.ab = 0.0f;
        L.prev_delta = 1;
        L.last_pdf = 0.0f;
        L.prev_sumW = 0.0f;
        L.prev_M = 0;
        L.pbr = L.pbg = L.pbb = L.pbcos = 0.0f;
        L.hit = 0;
        L.max_hits = P.bounces;
        L.clamp_neg = true;
        if (L.max_hits <= 0) return kPathEnd;
        set_ray(L, kExt, L.po, L.d, 1e30f);
        return kTrace;
    }

    // One ReSTIR indirect sample: BSDF sample at the primary hit, then a path
    // from bounce 1 (restir_shade, pathtrace.metal:1137-1174).
    __device__ int begin_restir_sample(Lane& L) {
        const float4 g0 = gb.p0[L.pixel], g1 = gb.p1[L.pixel], g2 = gb.p2[L.pixel], g3 = gb.p3[L.pixel];
        const V3 gpos = xyz(g0), gnor = xyz(g1), galb = xyz(g2), gwo = xyz(g3);
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
        float nx2, ny2, nz2;
        int pdelta = 0;
        if (pt.urand() < spec_prob) {
            float hx, hy, hz;
            pt.sample_ggx_half(alpha2, gnor.x, gnor.y, gnor.z, hx, hy, hz);
            const float md = gwo.x * hx + (gwo.y * hy + gwo.z * hz);
            if (!(md > 1e-8f)) return kPathEnd;
            nx2 = 2.0f * md * hx - gwo.x; ny2 = 2.0f * md * hy - gwo.y; nz2 = 2.0f * md * hz - gwo.z;
            pdelta = (rough <= 0.0201f) ? 1 : 0;
        } else {
            pt.sample_cosine(gnor.x, gnor.y, gnor.z, nx2, ny2, nz2);
            pdelta = 0;
        }
        const float ncos = nx2 * gnor.x + (ny2 * gnor.y + nz2 * gnor.z);
        if (!(ncos > 1e-8f)) return kPathEnd;
        {
            const float dn = 1.0f / sqrtf(nx2 * nx2 + ny2 * ny2 + nz2 * nz2);
            nx2 *= dn; ny2 *= dn; nz2 *= dn;
        }
        const Bsdf e0 = Tracer<S>::eval_bsdf(nx2, ny2, nz2, gwo.x, gwo.y, gwo.z, gnor.x, gnor.y, gnor.z, alpha2,
                                             metallic, spec_prob, galb.x, galb.y, galb.z);
        if (!(e0.pdf > 1e-12f)) return kPathEnd;
        const float cos0 = nx2 * gnor.x + (ny2 * gnor.y + nz2 * gnor.z);
        L.tr = e0.r * cos0 / e0.pdf;
        L.tg = e0.g * cos0 / e0.pdf;
        L.tb = e0.b * cos0 / e0.pdf;
        L.ar = L.ag = L.ab = 0.0f;
        L.po = V3{gpos.x + gnor.x * 0.001f, gpos.y + gnor.y * 0.001f, gpos.z + gnor.z * 0.001f};
        L.d = V3{nx2, ny2, nz2};
        L.last_pdf = e0.pdf;
        L.pbr = e0.r; L.pbg = e0.g; L.pbb = e0.b; L.pbcos = cos0;
        L.prev_sumW = L.m_sumWu;
        L.prev_M = L.m_M;
        L.prev_delta = pdelta;
        L.hit = 0;
        L.max_hits = P.bounces - 1;
        L.clamp_neg = false;
        if (L.max_hits <= 0) return kPathEnd;
        set_ray(L, kExt, L.po, L.d, 1e30f);
        return kTrace;
    }

    __device__ V3 clamped(float r, float g, float b) const {
        const float c = P.clamp_value;
        return c > 0.0f ? V3{fminf(r, c), fminf(g, c), fminf(b, c)} : V3{r, g, b};
    }

    __device__ void finish_pixel(Lane& L) {
        if (!SHADE) {
            const float mr = L.sr / float(L.n), mg = L.sg / float(L.n), mb = L.sb / float(L.n);
            rgb[3 * L.pixel] = tonemap8(mr);
            rgb[3 * L.pixel + 1] = tonemap8(mg);
            rgb[3 * L.pixel + 2] = tonemap8(mb);
            if (hdr) hdr[L.pixel] = make_float4(mr, mg, mb, 0.0f);
        } else {
            const float fn = float(max(L.n, 1));
            V3 c{L.sr / fn, L.sg / fn, L.sb / fn};
            if (L.kind == 3) c = clamped(c.x, c.y, c.z);
            hdr[L.pixel] = make_float4(c.x, c.y, c.z, 0.0f);
            const float n = float(L.n);
            write_var(L.pixel, L.n > 1 ? fmaxf(0.0f, (L.sl2 - L.sl * L.sl / n) / (n - 1.0f)) : 0.0f);
        }
        L.phase = kIdle;
    }

    // Accounts the path that just ended and starts the next sample; loops over
    // samples that finish without a ray. Returns kTrace or kPixelDone.
    __device__ int next_sample(Lane& L) {
        for (;;) {
            bool stop = false;
            if (L.kind == 1) {
                const V3 full = clamped(L.dr + L.ar, L.dg + L.ag, L.db + L.ab);
                L.sr += full.x; L.sg += full.y; L.sb += full.z;
                if (SHADE && lum_var) add_lum(L, full);
            } else {
                const V3 c = clamped(L.ar, L.ag, L.ab);
                L.sr += c.x; L.sg += c.y; L.sb += c.z;
                if (SHADE && lum_var) add_lum(L, c);
                if (!SHADE && P.adapt_min > 0) {
                    const float lum = c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f;
                    L.sl += lum;
                    L.sl2 += lum * lum;
                    const int n = L.n + 1;
                    if (n >= P.adapt_min && n > 1 && n % P.adapt_step == 0) {
                        const float mean = L.sl / float(n);
                        const float var = fmaxf(0.0f, (L.sl2 - L.sl * mean) / float(n - 1));
                        const float th = P.adapt_rel * mean + P.adapt_abs;
                        stop = var < th * th;
                    }
                }
            }
            ++L.n;
            ++L.s;
            if (stop || L.s >= P.spp) {
                finish_pixel(L);
                return kPixelDone;
            }
            const int r = L.kind == 1 ? begin_restir_sample(L) : begin_camera_sample(L);
            if (r == kTrace) return kTrace;
            L.ar = L.ag = L.ab = 0.0f;  // a sample without a path adds the direct term only
        }
    }

    // Starts pixel `px`. Returns kTrace, or kPixelDone when it needs no ray.
    __device__ int start_pixel(Lane& L, int px) {
        L.pixel = px;
        L.s = L.n = 0;
        L.sr = L.sg = L.sb = L.sl = L.sl2 = 0.0f;
        const uint32_t x = uint32_t(px) % uint32_t(P.width), y = uint32_t(px) / uint32_t(P.width);
        if (!SHADE) {
            L.kind = 3;
            pt.rng.seed(pixel_seed(P.seed, x, y));
        } else {
            const float4 g0 = gb.p0[px];
            L.kind = ibits(g0.w);
            if (L.kind == 0) {
                hdr[px] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                write_var(px, 0.0f);
                L.phase = kIdle;
                return kPixelDone;
            }
            pt.rng.seed(P.seed ^ (uint64_t(px + 1) * 0xBF58476D1CE4E5B9ULL) ^ 0x94D049BB133111EBULL);
            if (L.kind == 1) {
                if (!prepare_direct(L, px)) return kPixelDone;
            }
        }
        const int r = L.kind == 1 ? begin_restir_sample(L) : begin_camera_sample(L);
        if (r == kTrace) return kTrace;
        L.ar = L.ag = L.ab = 0.0f;
        return next_sample(L);
    }

    // Primary direct lighting from the merged reservoir (no shadow ray: the
    // visibility was reused), pathtrace.metal:1086-1134. Returns false when the
    // pixel is complete without indirect samples.
    __device__ bool prepare_direct(Lane& L, int px) {
        const float4 g0 = gb.p0[px], g1 = gb.p1[px], g2 = gb.p2[px], g3 = gb.p3[px], g4 = gb.p4[px];
        const V3 gpos = xyz(g0), gnor = xyz(g1), galb = xyz(g2), gwo = xyz(g3), gemi = xyz(g4);
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
        V3 prim{0.0f, 0.0f, 0.0f};
        L.m_sumWu = 0.0f;
        L.m_M = 0;
        const float4 m0 = rm.r0[px], m1 = rm.r1[px], m2 = rm.r2[px], m3 = rm.r3[px];
        const uint32_t valid = __float_as_uint(m3.z), M = __float_as_uint(m3.y);
        if (valid != 0u && M > 0u && m3.x > 0.0f && m2.w > 0.0f && m0.w > 0.0f) {
            const V3 lp = xyz(m0), ln = xyz(m1), le = xyz(m2);
            const float lalpha = m1.w;
            L.m_sumWu = m3.x;
            L.m_M = M;
            const float m_sumWs = m2.w;
            const float ldx = lp.x - gpos.x, ldy = lp.y - gpos.y, ldz = lp.z - gpos.z;
            const float dist = sqrtf(ldx * ldx + ldy * ldy + ldz * ldz);
            if (dist > 0.003f) {
                const float inv = 1.0f / dist;
                const float dxn = ldx * inv, dyn = ldy * inv, dzn = ldz * inv;
                const float lcos = -(ln.x * dxn + (ln.y * dyn + ln.z * dzn));
                const float scos = gnor.x * dxn + (gnor.y * dyn + gnor.z * dzn);
                if (lcos > 1e-8f && scos > 1e-8f) {
                    const Bsdf e = Tracer<S>::eval_bsdf(dxn, dyn, dzn, gwo.x, gwo.y, gwo.z, gnor.x, gnor.y, gnor.z,
                                                        alpha2, metallic, spec_prob, galb.x, galb.y, galb.z);
                    const float Fr = le.x * e.r * scos * lalpha, Fg = le.y * e.g * scos * lalpha,
                                Fb = le.z * e.b * scos * lalpha;
                    const float ph = lum3(Fr, Fg, Fb);
                    if (ph > 0.0f) {
                        const float pdf_ris = float(L.m_M) * ph / L.m_sumWu;
                        float wnee = 1.0f;
                        if (!P.nomis && e.pdf > 1e-20f) wnee = pdf_ris / (pdf_ris + e.pdf);
                        const float W = m_sumWs / (float(L.m_M) * ph);
                        prim = V3{Fr * W * wnee, Fg * W * wnee, Fb * W * wnee};
                    }
                }
            }
        }
        L.dr = gemi.x + prim.x;
        L.dg = gemi.y + prim.y;
        L.db = gemi.z + prim.z;
        if (P.bounces <= 1 || P.spp <= 0) {
            const V3 c = clamped(L.dr, L.dg, L.db);
            hdr[px] = make_float4(c.x, c.y, c.z, 0.0f);
            write_var(px, 0.0f);
            L.phase = kIdle;
            return false;
        }
        return true;
    }

    // End of one bounce (pathtrace.metal:708-725, 1244-1250).
    __device__ int post_bounce(Lane& L) {
        if (L.clamp_neg) {
            if (L.tr < 0.0f) L.tr = 0.0f;
            if (L.tg < 0.0f) L.tg = 0.0f;
            if (L.tb < 0.0f) L.tb = 0.0f;
        }
        float mx = fmaxf(L.tr, L.tg);
        mx = fmaxf(mx, L.tb);
        if (!(mx >= 1e-12f)) return kPathEnd;
        if (L.hit >= 1) {
            float rr = mx;
            if (rr < 0.05f) rr = 0.05f;
            if (rr > 0.9f) rr = 0.9f;
            if (!(pt.urand() <= rr)) return kPathEnd;
            const float ir = 1.0f / rr;
            L.tr *= ir; L.tg *= ir; L.tb *= ir;
        }
        ++L.hit;
        if (L.hit >= L.max_hits) return kPathEnd;
        set_ray(L, kExt, L.po, L.d, 1e30f);
        return kTrace;
    }

    // After the shadow ray of an opaque vertex: add the light sample, sample
    // the BSDF for the next direction (pathtrace.metal:666-707).
    __device__ int after_nee(Lane& L, bool lit) {
        const float dlr = lit ? L.nr : 0.0f, dlg = lit ? L.ng : 0.0f, dlb = lit ? L.nb : 0.0f;
        L.ar += dlr * L.tr;
        L.ag += dlg * L.tg;
        L.ab += dlb * L.tb;
        return continue_path(L);
    }
    // The light term of a kDual / kShadowEnd lane once its shadow ray is done.
    __device__ void add_light(Lane& L, bool lit) {
        L.ar += lit ? L.nr : 0.0f;
        L.ag += lit ? L.ng : 0.0f;
        L.ab += lit ? L.nb : 0.0f;
    }
    __device__ V3 shadow_origin(const Lane& L) const {
        return V3{L.p.x + L.pn.x * 0.002f, L.p.y + L.pn.y * 0.002f, L.p.z + L.pn.z * 0.002f};
    }
    // BSDF sample at the vertex kept in L and the end of the bounce.
    __device__ int continue_path(Lane& L) {
        const float wox = -L.d.x, woy = -L.d.y, woz = -L.d.z;
        const V3 pn = L.pn;
        float nx2, ny2, nz2;
        if (pt.urand() < L.h_spec) {
            float hx, hy, hz;
            pt.sample_ggx_half(L.h_alpha2, pn.x, pn.y, pn.z, hx, hy, hz);
            const float md = wox * hx + (woy * hy + woz * hz);
            if (!(md > 1e-8f)) return kPathEnd;
            nx2 = 2.0f * md * hx - wox; ny2 = 2.0f * md * hy - woy; nz2 = 2.0f * md * hz - woz;
            L.prev_delta = (L.h_rough <= 0.0201f) ? 1 : 0;
        } else {
            pt.sample_cosine(pn.x, pn.y, pn.z, nx2, ny2, nz2);
            L.prev_delta = 0;
        }
        const float ncos = nx2 * pn.x + (ny2 * pn.y + nz2 * pn.z);
        if (!(ncos > 1e-8f)) return kPathEnd;
        {
            const float dn = 1.0f / sqrtf(nx2 * nx2 + ny2 * ny2 + nz2 * nz2);
            nx2 *= dn; ny2 *= dn; nz2 *= dn;
        }
        const Bsdf e = Tracer<S>::eval_bsdf(nx2, ny2, nz2, wox, woy, woz, pn.x, pn.y, pn.z, L.h_alpha2, L.h_met,
                                            L.h_spec, L.h_ar, L.h_ag, L.h_ab);
        if (!(e.pdf > 1e-12f)) return kPathEnd;
        L.last_pdf = e.pdf;
        const float cos_i = nx2 * pn.x + (ny2 * pn.y + nz2 * pn.z);
        L.tr *= e.r * cos_i / e.pdf;
        L.tg *= e.g * cos_i / e.pdf;
        L.tb *= e.b * cos_i / e.pdf;
        L.prev_sumW = L.nsumW;
        L.prev_M = L.nM;
        L.pbr = e.r; L.pbg = e.g; L.pbb = e.b; L.pbcos = cos_i;
        L.po = V3{L.p.x + pn.x * 0.001f, L.p.y + pn.y * 0.001f, L.p.z + pn.z * 0.001f};
        L.d = V3{nx2, ny2, nz2};
        return post_bounce(L);
    }

    // A path vertex (the body of the bounce loops, pathtrace.metal:567-707).
    __device__ int shade_hit(Lane& L, const Surface& hit) {
        float pnx = hit.n_interp.x, pny = hit.n_interp.y, pnz = hit.n_interp.z;
        const float geom = hit.n_geom.x * pnx + (hit.n_geom.y * pny + hit.n_geom.z * pnz);
        if (geom < 0.0f) { pnx = -pnx; pny = -pny; pnz = -pnz; }
        const float ndot = L.d.x * pnx + (L.d.y * pny + L.d.z * pnz);
        bool front = true;
        if (ndot > 0.0f) { pnx = -pnx; pny = -pny; pnz = -pnz; front = false; }
        const float wox = -L.d.x, woy = -L.d.y, woz = -L.d.z;
        float rough = hit.rough;
        if (rough < 0.0f) rough = 0.0f;
        if (rough > 1.0f) rough = 1.0f;
        if (rough < 0.02f) rough = 0.02f;
        const float alpha = rough * rough;
        const float alpha2 = alpha * alpha;
        float metallic = hit.metallic;
        if (metallic < 0.0f) metallic = 0.0f;
        if (metallic > 1.0f) metallic = 1.0f;
        if (L.prev_delta == 1) {
            L.ar += hit.er * L.tr; L.ag += hit.eg * L.tg; L.ab += hit.eb * L.tb;
        } else {
            const float w = pt.emission_mis(hit, L.po.x, L.po.y, L.po.z, L.last_pdf, L.prev_sumW, L.prev_M, L.pbr,
                                            L.pbg, L.pbb, L.pbcos);
            L.ar += hit.er * L.tr * w; L.ag += hit.eg * L.tg * w; L.ab += hit.eb * L.tb * w;
        }
        if (hit.ior > 1.0001f) {
            const DielectricSample ds = pt.sample_dielectric(wox, woy, woz, pnx, pny, pnz, L.d.x, L.d.y, L.d.z,
                                                             hit.ior, front, rough, alpha2);
            if (!ds.valid) return kPathEnd;
            L.tr *= ds.eta_scale; L.tg *= ds.eta_scale; L.tb *= ds.eta_scale;
            float sgn = 1.0f;
            if (ds.n_x * pnx + (ds.n_y * pny + ds.n_z * pnz) < 0.0f) sgn = -1.0f;
            L.po = V3{hit.p.x + pnx * 0.001f * sgn, hit.p.y + pny * 0.001f * sgn, hit.p.z + pnz * 0.001f * sgn};
            L.d = V3{ds.n_x, ds.n_y, ds.n_z};
            L.prev_delta = 1;
            L.prev_sumW = 0.0f;
            L.prev_M = 0;
            return post_bounce(L);
        }
        float spec_prob = 0.05f + 0.95f * metallic;
        if (spec_prob > 0.98f) spec_prob = 0.98f;
        if (metallic >= 0.999f) spec_prob = 1.0f;
        L.p = hit.p;
        L.pn = V3{pnx, pny, pnz};
        L.h_ar = hit.ar; L.h_ag = hit.ag; L.h_ab = hit.ab;
        L.h_rough = rough; L.h_alpha2 = alpha2; L.h_met = metallic; L.h_spec = spec_prob;
        if (defer_nee) {
            L.phase = kNee;
            return kWait;
        }
        return nee(L);
    }

    // Light sampling at the vertex kept in L (the shading point, its normal,
    // the material, and wo = -L.d), then its shadow ray.
    __device__ int nee(Lane& L) {
        return finish_nee(L, pt.nee_prepare(L.p.x, L.p.y, L.p.z, L.pn.x, L.pn.y, L.pn.z, -L.d.x, -L.d.y, -L.d.z,
                                            L.h_rough, L.h_met, L.h_spec, L.h_ar, L.h_ag, L.h_ab));
    }
    __device__ int finish_nee(Lane& L, const typename Tracer<S>::NeeSample& ns) {
        L.nr = ns.r; L.ng = ns.g; L.nb = ns.b;
        L.nsumW = ns.sumW;
        L.nM = ns.M;
        if (ns.shadow) {
            if (dual_shadow) {
                L.nr *= L.tr;  // the products after_nee would add (dl * throughput of this vertex)
                L.ng *= L.tg;
                L.nb *= L.tb;
                L.sd = ns.d;
                L.sdist = ns.dist;
                const int r = continue_path(L);
                if (r == kTrace) {
                    L.phase = kDual;
                    return kTrace;
                }
                // no continuation: the shadow ray alone, then the path ends
                set_ray(L, kShadowEnd, ns.o, ns.d, ns.dist);
                return kTrace;
            }
            set_ray(L, kShadow, ns.o, ns.d, ns.dist);
            return kTrace;
        }
        return after_nee(L, false);
    }

    // Consumes the result of the ray the lane traced.
    __device__ int on_trace(Lane& L, Hit h) {
        if (!pt.sc.opaque && h.tri >= 0) {
            // Stochastic coverage (cast_alpha_ray, pathtrace.metal:536-550).
            const float opacity = pt.tri_alpha(h.tri);
            if (!(opacity >= 1.0f || (opacity > 0.0f && pt.urand() < opacity))) {
                L.traveled += h.t + 0.0001f;
                if (L.traveled < L.seq_dist) {
                    L.ro = L.base + L.rd * L.traveled;
                    L.tmax = L.seq_dist - L.traveled;
                    return kTrace;
                }
                h.tri = -1;
            }
        }
        if (L.phase == kShadow) return after_nee(L, h.tri < 0);
        if (L.phase == kShadowEnd) {
            add_light(L, h.tri < 0);
            return kPathEnd;
        }
        if (h.tri < 0) return kPathEnd;
        return shade_hit(L, pt.surface_at(h, L.ro, L.rd));
    }

    // Engine interface of k_persistent: start a pixel / consume a trace.
    // Both return kTrace (a ray is set up) or kPixelDone.
    __device__ int start(Lane& L, int px) { return start_pixel(L, px); }
    __device__ int step(Lane& L, const Hit& h) {
        const int r = on_trace(L, h);
        return r == kPathEnd ? next_sample(L) : r;
    }
    // Runs the deferred light sampling of a lane in phase kNee.
    __device__ int run_nee(Lane& L) {
        const int r = nee(L);
        return r == kPathEnd ? next_sample(L) : r;
    }
    // The same for the waiting lanes W of the wave together (Tracer::
    // nee_prepare_coop); every lane of the wave calls it. Returns the waiting
    // lane's step result (kTrace for the others).
    __device__ int run_nee_batch(Lane& L, unsigned long long W) {
        const typename Tracer<S>::NeeSample ns =
            pt.nee_prepare_coop(W, L.p.x, L.p.y, L.p.z, L.pn.x, L.pn.y, L.pn.z, -L.d.x, -L.d.y, -L.d.z, L.h_rough,
                                L.h_met, L.h_spec, L.h_ar, L.h_ag, L.h_ab);
        if (L.phase != kNee) return kTrace;
        const int r = finish_nee(L, ns);
        return r == kPathEnd ? next_sample(L) : r;
    }
    __device__ bool any_hit(const Lane& L) const {
        return (L.phase == kShadow || L.phase == kShadowEnd) && pt.sc.opaque;
    }
};


// restir_generate (pathtrace.metal:817-917) as a state machine: the primary
// ray (with its stochastic-coverage walk), then up to Mc light candidates,
// each with a shadow ray when its target function is positive.
struct GenLane {
    int phase;
    V3 ro, rd, base;
    float tmax, traveled, seq_dist;
    int pixel, j, Mc;
    V3 p, pn, wo, alb;
    float alpha2, met, spec;
    float sumWs, sumWu, sel_ph;
    bool have;
    // A candidate is kept as its light and folded barycentrics; the light
    // point is recomputed from them with the same arithmetic when needed.
    int sel_light;
    float sel_lu, sel_lv;
    int c_light;  // candidate waiting for its shadow ray
    float c_lu, c_lv, c_wu, c_ph;
};

// The path engine at 3 blocks per CU (with some spilling) for large trees,
// whose node fetches miss the caches more and want more waves to hide them
// (2.1 M triangles, 1.03 M nodes: shade 7.5 ms against 8.6 ms at 2 blocks).
// launch_wave takes it for trees of more than kDeepTreeNodes nodes; 3 blocks'
// stacks fit the LDS up to S = 48.
constexpr int kDeepTreeNodes = 1 << 19;

template <int S, bool SHADE>
struct PathEngineDeep : PathEngine<S, SHADE> {
    static_assert(S <= 48, "LDS holds 3 blocks up to S = 48");
    using PathEngine<S, SHADE>::PathEngine;
    static constexpr int kMinBlocks = 3;
};

template <int S>
struct GenEngine {
    using Lane = GenLane;
    static constexpr int kPass = kCntGenerate;
    // shadow rays: latency hiding wins over the spills (LDS allows 4 blocks at S = 32)
    static constexpr int kMinBlocks = S <= 32 ? 4 : S <= 48 ? 3 : 2;
    static constexpr bool kDeferNee = false;
    static constexpr bool kDualShadow = false;
    static constexpr bool kFused = true;  // traversal inside the persistent loop, with refill
    // Leaf batches in that loop: when half the traversing lanes wait or
    // after 5 iterations (C1 generate 3.84-3.87 ms against 3.97-4.02 at 3;
    // 6/8 with 5: 3.93, 4/8 with 8: 3.84-3.89; the call-based traversal of
    // the other passes keeps 3).
    static constexpr int kLeafMaxAge = 5;
    static constexpr int kLeafFrac8 = 4;
    Tracer<S>& pt;
    const FrameParams& P;
    const GBufferSoA& gb;
    const ReservoirSoA& ri;
    __device__ GenEngine(Tracer<S>& t, const FrameParams& p, const PassIO& io) : pt(t), P(p), gb(io.gb), ri(io.res) {}

    __device__ void set_ray(Lane& L, int phase, V3 o, V3 dir, float dist) {
        L.phase = phase;
        L.ro = o;
        L.base = o;
        L.rd = dir;
        L.tmax = dist;
        L.traveled = 0.0f;
        L.seq_dist = dist;
    }

    __device__ int start(Lane& L, int px) {
        const uint32_t x = uint32_t(px) % uint32_t(P.width), y = uint32_t(px) / uint32_t(P.width);
        L.pixel = px;
        pt.rng.seed(pixel_seed(P.seed, x, y));
        float sx, sy;
        pt.jitter(x, y, sx, sy);
        set_ray(L, kExt, V3{P.cam.x, P.cam.y, P.cam.z}, pt.primary_dir(sx, sy, true), 1e30f);
        return kTrace;
    }

    // A pixel's reservoir is written once, when the pixel is done: an early
    // "invalid" store at its start and the final one much later would reach
    // HBM twice (the line is evicted in between).
    __device__ void invalid_reservoir(int px) {
        const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        ri.r0[px] = zero; ri.r1[px] = zero; ri.r2[px] = zero; ri.r3[px] = zero;  // valid = 0
    }

    __device__ void finish_candidate(Lane& L, float wu, float ws, float ph, int li, float lu, float lv) {
        L.sumWu += wu;
        L.sumWs += ws;
        const float r = pt.urand();
        if (ws > 0.0f && r * L.sumWs < ws) {
            L.have = true;
            L.sel_light = li;
            L.sel_lu = lu;
            L.sel_lv = lv;
            L.sel_ph = ph;
        }
        ++L.j;
    }

    // Uniform point of light `li` at folded barycentrics (lu, lv).
    __device__ V3 light_point(const typename Tracer<S>::LightRec& Lt, float lu, float lv) const {
        const float4 l0 = Lt.a, l1 = Lt.b, l2 = Lt.c;
        const float lw = 1.0f - lu - lv;
        return V3{lw * l0.x + (lu * l0.w + lv * l1.z), lw * l0.y + (lu * l1.x + lv * l1.w),
                  lw * l0.z + (lu * l1.y + lv * l2.x)};
    }

    // Draws candidates until one needs a shadow ray (kTrace) or all are done.
    __device__ int next_candidate(Lane& L) {
        const DevScene& sc = pt.sc;
        while (L.j < L.Mc) {
            const float pick = pt.urand() * sc.light_total;
            const int li = pt.pick_light(pick);
            const typename Tracer<S>::LightRec Lt = pt.light(li);
            float lu = pt.urand(), lv = pt.urand();
            if (lu + lv > 1.0f) { lu = 1.0f - lu; lv = 1.0f - lv; }
            const V3 q = light_point(Lt, lu, lv);
            const float ldx = q.x - L.p.x, ldy = q.y - L.p.y, ldz = q.z - L.p.z;
            const float dist = sqrtf(ldx * ldx + ldy * ldy + ldz * ldz);
            if (dist > 0.003f) {
                const float4 l2 = Lt.c, l3 = Lt.d;
                const float inv = 1.0f / dist;
                const float dxn = ldx * inv, dyn = ldy * inv, dzn = ldz * inv;
                const float lcos = -(l2.y * dxn + (l2.z * dyn + l2.w * dzn));
                const float scos = L.pn.x * dxn + (L.pn.y * dyn + L.pn.z * dzn);
                if (lcos > 1e-8f && scos > 1e-8f) {
                    float ll = lum3(l3.x, l3.y, l3.z);
                    if (ll < 1e-6f) ll = 1e-6f;
                    const float qd = dist * dist * ll / (lcos * sc.light_total);
                    if (qd > 1e-12f) {
                        const Bsdf e = Tracer<S>::eval_bsdf(dxn, dyn, dzn, L.wo.x, L.wo.y, L.wo.z, L.pn.x, L.pn.y,
                                                            L.pn.z, L.alpha2, L.met, L.spec, L.alb.x, L.alb.y,
                                                            L.alb.z);
                        const float ph = lum3(l3.x * e.r * scos * l3.w, l3.y * e.g * scos * l3.w,
                                              l3.z * e.b * scos * l3.w);
                        if (ph > 0.0f) {
                            L.c_wu = ph / qd;
                            L.c_ph = ph;
                            L.c_light = li;
                            L.c_lu = lu;
                            L.c_lv = lv;
                            const V3 so{L.p.x + L.pn.x * 0.002f, L.p.y + L.pn.y * 0.002f, L.p.z + L.pn.z * 0.002f};
                            set_ray(L, kShadow, so, V3{dxn, dyn, dzn}, dist - 0.003f);
                            return kTrace;
                        }
                        finish_candidate(L, 0.0f, 0.0f, ph, li, lu, lv);
                        continue;
                    }
                }
            }
            finish_candidate(L, 0.0f, 0.0f, 0.0f, li, lu, lv);
        }
        if (L.have && L.sumWs > 0.0f && L.sumWu > 0.0f && L.sel_ph > 0.0f) {
            const int px = L.pixel;
            const typename Tracer<S>::LightRec Lt = pt.light(L.sel_light);
            const float4 l2 = Lt.c, l3 = Lt.d;
            const V3 q = light_point(Lt, L.sel_lu, L.sel_lv);
            ri.r0[px] = make_float4(q.x, q.y, q.z, L.sel_ph);
            ri.r1[px] = make_float4(l2.y, l2.z, l2.w, l3.w);
            ri.r2[px] = make_float4(l3.x, l3.y, l3.z, L.sumWs);
            ri.r3[px] = make_float4(L.sumWu, __uint_as_float(uint32_t(L.Mc)), __uint_as_float(1u), 0.0f);
        } else {
            invalid_reservoir(L.pixel);
        }
        return kPixelDone;
    }

    __device__ int primary_hit(Lane& L, const Surface& hit) {
        const int px = L.pixel;
        const V3 rd = L.rd;
        float pnx = hit.n_interp.x, pny = hit.n_interp.y, pnz = hit.n_interp.z;
        const float geom = hit.n_geom.x * pnx + (hit.n_geom.y * pny + hit.n_geom.z * pnz);
        if (geom < 0.0f) { pnx = -pnx; pny = -pny; pnz = -pnz; }
        const float ndot = rd.x * pnx + (rd.y * pny + rd.z * pnz);
        if (ndot > 0.0f) { pnx = -pnx; pny = -pny; pnz = -pnz; }
        const float wox = -rd.x, woy = -rd.y, woz = -rd.z;
        float rough = hit.rough;
        if (rough < 0.0f) rough = 0.0f;
        if (rough > 1.0f) rough = 1.0f;
        float metallic = hit.metallic;
        if (metallic < 0.0f) metallic = 0.0f;
        if (metallic > 1.0f) metallic = 1.0f;
        const bool dielectric = hit.ior > 1.0001f;
        gb.p0[px] = make_float4(hit.p.x, hit.p.y, hit.p.z, __int_as_float(dielectric ? 3 : 1));
        gb.p1[px] = make_float4(pnx, pny, pnz, rough);
        gb.p2[px] = make_float4(hit.ar, hit.ag, hit.ab, metallic);
        gb.p3[px] = make_float4(wox, woy, woz, dielectric ? hit.ior : 1.0f);
        gb.p4[px] = make_float4(hit.er, hit.eg, hit.eb, 0.0f);
        if (dielectric || pt.sc.nlights == 0 || !(pt.sc.light_total > 1e-12f)) {
            invalid_reservoir(px);
            return kPixelDone;
        }
        int Mc = P.restir_candidates;
        if (Mc < 1) Mc = 1;
        if (Mc > 32) Mc = 32;
        float r2 = rough;
        if (r2 < 0.02f) r2 = 0.02f;
        float spec_prob = 0.05f + 0.95f * metallic;
        if (spec_prob > 0.98f) spec_prob = 0.98f;
        if (metallic >= 0.999f) spec_prob = 1.0f;
        L.Mc = Mc;
        L.j = 0;
        L.p = hit.p;
        L.pn = V3{pnx, pny, pnz};
        L.wo = V3{wox, woy, woz};
        L.alb = V3{hit.ar, hit.ag, hit.ab};
        L.alpha2 = (r2 * r2) * (r2 * r2);
        L.met = metallic;
        L.spec = spec_prob;
        L.sumWs = L.sumWu = L.sel_ph = 0.0f;
        L.have = false;
        return next_candidate(L);
    }

    __device__ int step(Lane& L, Hit h) {
        if (!pt.sc.opaque && h.tri >= 0) {
            const float opacity = pt.tri_alpha(h.tri);
            if (!(opacity >= 1.0f || (opacity > 0.0f && pt.urand() < opacity))) {
                L.traveled += h.t + 0.0001f;
                if (L.traveled < L.seq_dist) {
                    L.ro = L.base + L.rd * L.traveled;
                    L.tmax = L.seq_dist - L.traveled;
                    return kTrace;
                }
                h.tri = -1;
            }
        }
        if (L.phase == kShadow) {
            finish_candidate(L, L.c_wu, h.tri < 0 ? L.c_wu : 0.0f, L.c_ph, L.c_light, L.c_lu, L.c_lv);
            return next_candidate(L);
        }
        if (h.tri < 0) {
            const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            const int px = L.pixel;
            gb.p0[px] = zero; gb.p1[px] = zero; gb.p2[px] = zero; gb.p3[px] = zero; gb.p4[px] = zero;
            invalid_reservoir(px);
            return kPixelDone;
        }
        return primary_hit(L, pt.surface_at(h, L.ro, L.rd));
    }
    __device__ bool any_hit(const Lane& L) const { return L.phase == kShadow && pt.sc.opaque; }
};

// Image bands of the work queues: band b holds rows [band_row(b), band_row(b + 1)).
__device__ __forceinline__ int band_row(int b, int height) { return (b * height + kBands - 1) / kBands; }

// Work of a persistent pass, split into kBands bands of image rows. A block
// starts on band blockIdx.x % kBands: workgroups are dispatched to the XCDs in
// turn, so each XCD's L2 serves the rays of one part of the image; a wave whose
// band is empty moves on to the next one. Band b's positions are taken from
// next[b]. Without `order`, position i of band b is pixel
// band_row(b) * width + i; with it, i < order_n[2b] reads order[s + i] and the
// others order[e - 1 - (i - order_n[2b])], for the band's segment [s, e) of
// `order` (order_n[2b] + order_n[2b + 1] positions).
struct BandQueue {
    unsigned int* next;
    const unsigned int* order;
    const unsigned int* order_n;
};

#ifndef RM_NEE_MAX_AGE
#define RM_NEE_MAX_AGE 2
#endif
#ifndef RM_NEE_FRAC8
#define RM_NEE_FRAC8 6
#endif
constexpr int kNeeMaxAge = RM_NEE_MAX_AGE;
#ifndef RM_TRAV_EXIT8
#define RM_TRAV_EXIT8 4
#endif
constexpr int kTravExit8 = RM_TRAV_EXIT8;
constexpr int kNeeFrac8 = RM_NEE_FRAC8;  // waiting share of the active lanes, in eighths, that runs the batch

// The persistent loop shared by the engines: every iteration traces one ray
// per active lane, then advances each lane; idle lanes take new pixels from
// the queue (one atomic per wave) until every band is exhausted.
template <int S, class E>
__global__ __launch_bounds__(kBlock, E::kMinBlocks) void k_persistent(DevScene sc, FrameParams P, PassIO io,
                                                                      BandQueue q, unsigned long long* cnt) {
    __shared__ uint32_t stack[lds_stack_entries<S>() * kBlock];
    Tracer<S> pt(sc, P, stack + threadIdx.x);
    pt.stage_lights();
    E eng(pt, P, io);
    if constexpr (E::kDeferNee) eng.defer_nee = true;
    if constexpr (E::kDualShadow) eng.dual_shadow = sc.opaque != 0;
    int nee_age = 0;  // iterations the wave's waiting lanes have waited
    const unsigned lane = __lane_id();
    int band = int(blockIdx.x) % kBands, tried = 0;
    unsigned seg0 = 0, seg1 = 0, front = 0, size = 0;
    auto enter = [&](int b) {
        seg0 = unsigned(band_row(b, P.height)) * unsigned(P.width);
        seg1 = unsigned(band_row(b + 1, P.height)) * unsigned(P.width);
        front = q.order ? q.order_n[2 * b] : 0u;
        size = q.order ? front + q.order_n[2 * b + 1] : seg1 - seg0;
    };
    enter(band);
    typename E::Lane L;
    L.phase = kIdle;
    bool exhausted = false;
    // E::kFused: the lane's ray is traversed in the loop below (trav) until
    // it is done (ready), then handed to the engine.
    typename Tracer<S>::Trav T;
    bool trav = false, ready = false;
    auto begin = [&]() {
        if (pt.trav_begin(T, L.ro, L.rd, L.tmax, eng.any_hit(L))) trav = true;
        else ready = true;
    };
    for (;;) {
        if (!exhausted) {
            const unsigned long long need = __ballot(L.phase == kIdle);
            if (need) {
                const int count = __popcll(need);
                const int leader = __ffsll((long long)need) - 1;
                unsigned base = 0;
                if (int(lane) == leader) base = atomicAdd(q.next + band, unsigned(count));
                base = __shfl(base, leader);
                if (L.phase == kIdle) {
                    const unsigned rank = __builtin_amdgcn_mbcnt_hi(unsigned(need >> 32),
                                                                    __builtin_amdgcn_mbcnt_lo(unsigned(need), 0u));
                    const unsigned i = base + rank;
                    if (i < size) {
                        const unsigned px = !q.order ? seg0 + i : i < front ? q.order[seg0 + i]
                                                                            : q.order[seg1 - 1u - (i - front)];
                        L.phase = kExt;  // start() sets the real phase
                        if (eng.start(L, int(px)) != kTrace) L.phase = kIdle;
                        else if (E::kFused) begin();
                    }
                }
                if (base + unsigned(count) >= size) {  // this band is done: the next one
                    if (++tried == kBands) {
                        exhausted = true;
                    } else {
                        band = band + 1 == kBands ? 0 : band + 1;
                        enter(band);
                    }
                }
            }
        }
        const unsigned long long act = __ballot(L.phase != kIdle);
        if (act == 0) {
            if (exhausted) break;
            continue;
        }
        if constexpr (E::kDeferNee) {
            // Lanes reaching a shading point wait until three quarters of
            // the active lanes do (which includes the case that no active lane
            // has a ray to trace) or until they have waited kNeeMaxAge
            // iterations; then the wave samples lights for all of them at once
            // (Tracer::nee_prepare_coop: ceil(Mc * waiting / 64) candidate
            // evaluations per lane and an Mc-step fold, against Mc evaluations
            // whenever any lane samples alone). A lane's own sequence of
            // operations and random draws is unchanged. With E::kTailNoWait,
            // once the queue is exhausted the wave's remaining lanes are the
            // long per-pixel chains that set the end of the pass, and they no
            // longer wait: C1 shade 10.4-10.5 -> 10.2-10.4 ms, C2 7.7 -> 7.3 ms
            // (the same for leaf batches made both slower).
            const unsigned long long wait = __ballot(L.phase == kNee);
            if (wait) {
                const int nw = __popcll(wait), na = __popcll(act);
                if (8 * nw >= kNeeFrac8 * na || ++nee_age >= kNeeMaxAge || (E::kTailNoWait && exhausted)) {
                    nee_age = 0;
                    const bool waiting = L.phase == kNee;
                    const int r = pt.ris_coop_applies() ? eng.run_nee_batch(L, wait)
                                                        : (waiting ? eng.run_nee(L) : int(kTrace));
                    if (waiting) {
                        if (r == kPixelDone) L.phase = kIdle;
                        else if (E::kFused && r == kTrace) begin();
                    }
                }
            }
        }
        if constexpr (E::kFused) {
            // Lanes whose ray is done take their step; a new ray starts at once.
            if (ready) {
                ready = false;
                const int r = eng.step(L, T.h);
                if (r == kTrace) begin();
                else if (r != kWait) L.phase = kIdle;
            }
            // Traverse until no more than kTravExit8 eighths of the working lanes
            // still traverse, so lanes with short rays take new ones instead of
            // waiting for the wave's longest traversal.
            {
                const int nwork = __popcll(__ballot(L.phase != kIdle));
                int leaf_age = 0;
                unsigned long long tv;
                do {
                    if (trav && !T.pend) pt.trav_node(T);
                    if (Tracer<S>::leaf_batch_now(__ballot(trav && T.pend), __ballot(trav), leaf_age,
                                                  E::kLeafMaxAge, E::kLeafFrac8) && trav &&
                        T.pend && pt.trav_leaves(T)) {
                        trav = false;
                        ready = true;
                    }
                    if (trav && T.last && !T.pend) {
                        trav = false;
                        ready = true;
                    }
                    tv = __ballot(trav);
                } while (8 * __popcll(tv) > kTravExit8 * nwork);
            }
        } else if constexpr (E::kDualShadow) {
            // A lane with both rays pending (kDual) hands its shadow ray to an
            // idle lane of the wave when there is one (the k-th such lane to
            // the k-th idle lane) and traces its extension ray meanwhile, so
            // the shadow ray leaves the pixel's chain of dependent rays; idle
            // lanes exist once the pixel queue is exhausted, in the tail that
            // sets the end of the pass. Without a helper the lane traces its
            // shadow ray first and the extension in the next iteration.
            const unsigned long long D = __ballot(L.phase == kDual), I = __ballot(L.phase == kIdle);
            int partner = -1;
            bool helper = false;
            if (D && I) {
                const int pairs = min(__popcll(D), __popcll(I));
                const unsigned long long below = (1ull << lane) - 1ull;
                if (L.phase == kDual) {
                    const int r = __popcll(D & below);
                    if (r < pairs) partner = nth_set_bit(I, r);
                } else if (L.phase == kIdle) {
                    const int r = __popcll(I & below);
                    if (r < pairs) {
                        partner = nth_set_bit(D, r);
                        helper = true;
                    }
                }
            }
            const int src = partner >= 0 ? partner : int(lane);
            const bool dual = L.phase == kDual;
            const V3 so = dual ? eng.shadow_origin(L) : V3{0.0f, 0.0f, 0.0f};
            const V3 bo{__shfl(so.x, src), __shfl(so.y, src), __shfl(so.z, src)};
            const V3 bd{__shfl(L.sd.x, src), __shfl(L.sd.y, src), __shfl(L.sd.z, src)};
            const float bdist = __shfl(L.sdist, src);
            V3 ro = L.ro, rd = L.rd;
            float tmax = L.tmax;
            bool any = eng.any_hit(L);
            if (helper || (dual && partner < 0)) {
                ro = bo;
                rd = bd;
                tmax = bdist;
                any = true;
            }
            const bool tracing = helper || dual || L.phase == kExt || L.phase == kShadow || L.phase == kShadowEnd;
            Hit h;
            h.t = 0.0f;
            h.tri = -1;
            h.u = h.v = 0.0f;
            if (tracing) h = pt.intersect(ro, rd, tmax, any);
            const int occluded = __shfl(int(h.tri >= 0), src);
            if (dual) {
                eng.add_light(L, (partner >= 0 ? occluded : int(h.tri >= 0)) == 0);
                L.phase = kExt;  // the extension ray stays in L.ro / L.rd
                if (partner >= 0) {
                    const int r = eng.step(L, h);
                    if (r != kTrace && r != kWait) L.phase = kIdle;
                }
            } else if (tracing && !helper) {
                const int r = eng.step(L, h);
                if (r != kTrace && r != kWait) L.phase = kIdle;
            }
        } else if (L.phase == kExt || L.phase == kShadow) {
            const Hit h = pt.intersect(L.ro, L.rd, L.tmax, eng.any_hit(L));
            const int r = eng.step(L, h);
            if (r != kTrace && r != kWait) L.phase = kIdle;
        }
    }
    pt.flush(cnt, E::kPass);
}

// Blocks of `kernel` resident on the current device at once (CUs x blocks per
// CU), looked up once per device and kernel.
template <int S>
int resident_blocks(const void* kernel) {
    static std::mutex mu;
    static std::map<std::pair<int, const void*>, int> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = -1;
    std::lock_guard<std::mutex> lock(mu);
    const auto key = std::make_pair(dev, kernel);
    const auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int cus = 1, occ = 0;
    if (dev >= 0) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
            cus = prop.multiProcessorCount;
    }
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, kBlock, 0) != hipSuccess || occ < 1) occ = 1;
    return cache[key] = cus * occ;
}

// Launches k_persistent<S, E> with one resident wave set.
template <int S, class E>
void launch_persistent(const DevScene& sc, const FrameParams& P, const PassIO& io, const BandQueue& q,
                       unsigned long long* cnt, hipStream_t st) {
    const int resident = resident_blocks<S>(reinterpret_cast<const void*>(&k_persistent<S, E>));
    const int npx = P.width * P.height;
    const int need = (npx + kBlock - 1) / kBlock;
    const dim3 grid(unsigned(need < resident ? need : resident)), block(kBlock);
    hipLaunchKernelGGL((k_persistent<S, E>), grid, block, 0, st, sc, P, io, q, cnt);
}

// Work order of restir_shade. Per-pixel chains differ widely in length, and a
// persistent lane that takes a long one late sets the end of the pass, so
// within each band's segment of `order` (the band's rows) pixels whose primary
// hit is a dielectric (camera paths, the longest chains) are queued first
// (from the segment's start), then the ReSTIR pixels (from its end). Pixels
// without a primary hit get their result here (black, zero variance, as
// start_pixel writes them) and are not queued. The order only changes which
// lane runs a pixel, never its result: every pixel's random sequence is
// seeded from its index.
__global__ __launch_bounds__(kBlock) void k_shade_order(FrameParams P, GBufferSoA gb, float4* __restrict__ hdr,
                                                        float* __restrict__ lum_var, unsigned* __restrict__ order,
                                                        unsigned* __restrict__ order_n) {
    const int npx = P.width * P.height;
    const int px = int(blockIdx.x) * kBlock + int(threadIdx.x);
    int kind = 0, band = 0;
    if (px < npx) {
        kind = ibits(gb.p0[px].w);
        if (kind == 0) {
            hdr[px] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (lum_var) lum_var[px] = 0.0f;
        }
        band = (px / P.width) * kBands / P.height;
    }
    const bool first = kind != 0 && kind != 1, last = kind == 1;
    bool pending = first || last;
    for (;;) {  // one round per band present in the wave (one or two)
        const unsigned long long m = __ballot(pending);
        if (!m) break;
        const int lb = __shfl(band, __ffsll((long long)m) - 1);
        const bool mine = pending && band == lb;
        const unsigned a = wave_append_pos(mine && first, order_n + 2 * lb);
        const unsigned c = wave_append_pos(mine && last, order_n + 2 * lb + 1);
        if (mine && first) order[unsigned(band_row(lb, P.height) * P.width) + a] = unsigned(px);
        if (mine && last) order[unsigned(band_row(lb + 1, P.height) * P.width) - 1u - c] = unsigned(px);
        if (mine) pending = false;
    }
}

// ------------------------------------------------------------------ wavefront
// The same engines split into kernels by role. Each pixel's lane state lives
// in HBM between rays, one 32-bit word per array of `stride` slots (word k of
// slot i at k * stride + i), in two buffers: a trace kernel at full occupancy
// reads only each slot's pending ray and writes its hit, then a logic kernel
// loads the whole state, advances the engine by one step (the code the
// persistent loop runs) and appends the slots that trace again, compacted
// with one atomic per wave, to the other buffer. Active slots only ever
// decrease, so the host steers with a count read back a few iterations late;
// below a tail threshold the persistent loop resumes the remaining slots from
// their saved state. Every slot runs exactly the per-pixel sequence of the
// persistent engine, so results are identical to it.
template <class Lane>
struct Slot {
    static_assert(sizeof(Lane) % 4 == 0, "lane state is stored as 32-bit words");
    static constexpr int kLaneWords = int(sizeof(Lane) / 4);
    static constexpr int kWords = kLaneWords + 1;  // + random-number state
    static constexpr int kPhase = int(__builtin_offsetof(Lane, phase) / 4);
    static constexpr int kRo = int(__builtin_offsetof(Lane, ro) / 4);
    static constexpr int kRd = int(__builtin_offsetof(Lane, rd) / 4);
    static constexpr int kTmax = int(__builtin_offsetof(Lane, tmax) / 4);

    static __device__ __forceinline__ void load(const uint32_t* __restrict__ st, size_t stride, size_t i, Lane& L,
                                                uint32_t& rng) {
        uint32_t w[kLaneWords];
#pragma unroll
        for (int k = 0; k < kLaneWords; ++k) w[k] = st[size_t(k) * stride + i];
        __builtin_memcpy(&L, w, sizeof(Lane));
        rng = st[size_t(kLaneWords) * stride + i];
    }
    static __device__ __forceinline__ void store(uint32_t* __restrict__ st, size_t stride, size_t i, const Lane& L,
                                                 uint32_t rng) {
        uint32_t w[kLaneWords];
        __builtin_memcpy(w, &L, sizeof(Lane));
#pragma unroll
        for (int k = 0; k < kLaneWords; ++k) st[size_t(k) * stride + i] = w[k];
        st[size_t(kLaneWords) * stride + i] = rng;
    }
};

// Control words of the wavefront loop.
enum : int { kWfCount0 = 0, kWfCount1 = 1, kWfFetchTrace = 2, kWfFetchLogic = 3, kWfCtlWords = 4 };

// Next 64 entries of [0, n) for the calling wave, from the fetch counter.
__device__ __forceinline__ unsigned wave_fetch(unsigned* fetch) {
    unsigned base = 0;
    if (__lane_id() == 0) base = atomicAdd(fetch, 64u);
    return __shfl(base, 0);
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Starts every pixel; pixels that need a ray are appended to buffer 0.
template <int S, class E>
__global__ __launch_bounds__(kBlock) void k_wf_start(DevScene sc, FrameParams P, PassIO io, uint32_t* __restrict__ dst,
                                                     size_t stride, unsigned* __restrict__ ctl) {
    using W = Slot<typename E::Lane>;
    Tracer<S> pt(sc, P, nullptr);
    E eng(pt, P, io);
    const int px = int(blockIdx.x) * kBlock + int(threadIdx.x);
    typename E::Lane L;
    bool keep = false;
    if (px < P.width * P.height) {
        L.phase = kExt;  // start() sets the real phase
        keep = eng.start(L, px) == kTrace;
    }
    const unsigned j = wave_append_pos(keep, ctl + kWfCount0);
    if (keep) W::store(dst, stride, j, L, pt.rng.state);
}

// One ray per active slot of buffer `cur`: hit (t, u, v, triangle) into hits.
// Clears the count the next logic kernel appends to and its fetch counter.
template <int S, class E>
__global__ __launch_bounds__(kBlock) void k_wf_trace(DevScene sc, FrameParams P, const uint32_t* __restrict__ src,
                                                     size_t stride, uint4* __restrict__ hits, unsigned* __restrict__ ctl,
                                                     int cur, unsigned long long* cnt) {
    using W = Slot<typename E::Lane>;
    __shared__ uint32_t stack[lds_stack_entries<S>() * kBlock];
    Tracer<S> pt(sc, P, stack + threadIdx.x);
    const unsigned n = ctl[kWfCount0 + cur];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl[kWfCount0 + (cur ^ 1)] = 0u;
        ctl[kWfFetchLogic] = 0u;
    }
    for (;;) {
        const unsigned base = wave_fetch(ctl + kWfFetchTrace);
        if (base >= n) break;
        const unsigned i = base + __lane_id();
        if (i < n) {
            const size_t k = i;
            const int phase = int(src[size_t(W::kPhase) * stride + k]);
            const V3 ro{__uint_as_float(src[size_t(W::kRo) * stride + k]),
                        __uint_as_float(src[size_t(W::kRo + 1) * stride + k]),
                        __uint_as_float(src[size_t(W::kRo + 2) * stride + k])};
            const V3 rd{__uint_as_float(src[size_t(W::kRd) * stride + k]),
                        __uint_as_float(src[size_t(W::kRd + 1) * stride + k]),
                        __uint_as_float(src[size_t(W::kRd + 2) * stride + k])};
            const float tmax = __uint_as_float(src[size_t(W::kTmax) * stride + k]);
            const Hit h = pt.intersect(ro, rd, tmax, phase == kShadow && sc.opaque);
            hits[k] = make_uint4(__float_as_uint(h.t), __float_as_uint(h.u), __float_as_uint(h.v), uint32_t(h.tri));
        }
    }
    if (cnt) {
        const uint32_t rays = wave_sum(pt.tl.rays), nodes = wave_sum(pt.tl.nodes), tris = wave_sum(pt.tl.tris);
        const unsigned long long over = __ballot(pt.tl.overflow != 0);
        if (__lane_id() == 0) {
            if (rays) atomicAdd(cnt + 3 * E::kPass, (unsigned long long)rays);
            if (nodes) atomicAdd(cnt + 3 * E::kPass + 1, (unsigned long long)nodes);
            if (tris) atomicAdd(cnt + 3 * E::kPass + 2, (unsigned long long)tris);
            if (over) atomicAdd(cnt + kCntOverflow, (unsigned long long)__popcll(over));
        }
    }
}

// Advances every slot of buffer `cur` by its traced hit; slots that trace
// again go to buffer cur ^ 1. Clears the trace fetch counter.
template <int S, class E>
__global__ __launch_bounds__(kBlock, E::kLogicBlocks) void k_wf_logic(DevScene sc, FrameParams P, PassIO io,
                                                                     const uint32_t* __restrict__ src,
                                                                     uint32_t* __restrict__ dst, size_t stride,
                                                                     const uint4* __restrict__ hits,
                                                                     unsigned* __restrict__ ctl, int cur) {
    using W = Slot<typename E::Lane>;
    Tracer<S> pt(sc, P, nullptr);
    E eng(pt, P, io);
    const unsigned n = ctl[kWfCount0 + cur];
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl[kWfFetchTrace] = 0u;
    for (;;) {
        const unsigned base = wave_fetch(ctl + kWfFetchLogic);
        if (base >= n) break;
        const unsigned i = base + __lane_id();
        typename E::Lane L;
        bool keep = false;
        if (i < n) {
            W::load(src, stride, i, L, pt.rng.state);
            const uint4 hw = hits[i];
            const Hit h{__uint_as_float(hw.x), __uint_as_float(hw.y), __uint_as_float(hw.z), int(hw.w)};
            keep = eng.step(L, h) == kTrace;
        }
        const unsigned j = wave_append_pos(keep, ctl + kWfCount0 + (cur ^ 1));
        if (keep) W::store(dst, stride, j, L, pt.rng.state);
    }
}

// The persistent loop over the slots left in buffer `cur`: lanes resume saved
// states (from the trace fetch counter) instead of starting pixels.
template <int S, class E>
__global__ __launch_bounds__(kBlock, E::kMinBlocks) void k_wf_finish(DevScene sc, FrameParams P, PassIO io,
                                                                     const uint32_t* __restrict__ src, size_t stride,
                                                                     unsigned* __restrict__ ctl, int cur,
                                                                     unsigned long long* cnt) {
    using W = Slot<typename E::Lane>;
    __shared__ uint32_t stack[lds_stack_entries<S>() * kBlock];
    Tracer<S> pt(sc, P, stack + threadIdx.x);
    E eng(pt, P, io);
    const unsigned n = ctl[kWfCount0 + cur];
    unsigned* fetch = ctl + kWfFetchTrace;
    const unsigned lane = __lane_id();
    typename E::Lane L;
    L.phase = kIdle;
    bool exhausted = false;
    for (;;) {
        if (!exhausted) {
            const unsigned long long need = __ballot(L.phase == kIdle);
            if (need) {
                const int count = __popcll(need);
                const int leader = __ffsll((long long)need) - 1;
                unsigned base = 0;
                if (int(lane) == leader) base = atomicAdd(fetch, unsigned(count));
                base = __shfl(base, leader);
                if (L.phase == kIdle) {
                    const unsigned i = base + __builtin_amdgcn_mbcnt_hi(unsigned(need >> 32),
                                                                        __builtin_amdgcn_mbcnt_lo(unsigned(need), 0u));
                    if (i < n) W::load(src, stride, i, L, pt.rng.state);
                }
                if (base + unsigned(count) >= n) exhausted = true;
            }
        }
        if (__ballot(L.phase != kIdle) == 0) {
            if (exhausted) break;
            continue;
        }
        if (L.phase != kIdle) {
            const Hit h = pt.intersect(L.ro, L.rd, L.tmax, eng.any_hit(L));
            if (eng.step(L, h) != kTrace) L.phase = kIdle;
        }
    }
    pt.flush(cnt, E::kPass);
}

#ifndef RM_WAVE_BATCH
#define RM_WAVE_BATCH 4  // iterations between count read-backs
#endif
#ifndef RM_WAVE_TAIL8
#define RM_WAVE_TAIL8 8  // tail threshold in eighths of the persistent loop's resident lanes
#endif
#ifndef RM_WAVE_MAX_ITERS
#define RM_WAVE_MAX_ITERS 4096
#endif

template <int S, class E>
int run_wavefront(const DevScene& sc, const FrameParams& P, const PassIO& io, const WaveBufs& wb,
                  unsigned long long* cnt, hipStream_t st, WaveRun* run) {
    const int trace_blocks = resident_blocks<S>(reinterpret_cast<const void*>(&k_wf_trace<S, E>));
    const int logic_blocks = resident_blocks<S>(reinterpret_cast<const void*>(&k_wf_logic<S, E>));
    const int finish_blocks = resident_blocks<S>(reinterpret_cast<const void*>(&k_wf_finish<S, E>));
    const unsigned npx = unsigned(P.width) * unsigned(P.height);
    if (size_t(npx) > wb.stride) return int(hipErrorInvalidValue);
    const size_t stride = wb.stride;
    hipError_t e = hipMemsetAsync(wb.ctl, 0, kWfCtlWords * sizeof(unsigned), st);
    if (e != hipSuccess) return int(e);
    const dim3 block(kBlock);
    hipLaunchKernelGGL((k_wf_start<S, E>), dim3((npx + kBlock - 1) / kBlock), block, 0, st, sc, P, io, wb.state[0],
                       stride, wb.ctl);
    // Grids: the resident set, or fewer blocks when the last known count
    // (an upper bound of the current one) needs fewer.
    unsigned known = npx;
    auto grid_for = [&](int resident) {
        const unsigned need = known > 0u ? (known + kBlock - 1u) / kBlock : 1u;
        return dim3(need < unsigned(resident) ? need : unsigned(resident));
    };
    const unsigned tail = wb.tail ? unsigned(size_t(finish_blocks) * kBlock * RM_WAVE_TAIL8 / 8) : 0u;
    int cur = 0, pending = -1, batch = 0, it = 0;
    while (it < RM_WAVE_MAX_ITERS) {
        for (int b = 0; b < RM_WAVE_BATCH; ++b, ++it) {
            hipLaunchKernelGGL((k_wf_trace<S, E>), grid_for(trace_blocks), block, 0, st, sc, P, wb.state[cur], stride,
                               wb.hits, wb.ctl, cur, cnt);
            hipLaunchKernelGGL((k_wf_logic<S, E>), grid_for(logic_blocks), block, 0, st, sc, P, io, wb.state[cur],
                               wb.state[cur ^ 1], stride, wb.hits, wb.ctl, cur);
            cur ^= 1;
        }
        const int slot = batch & 1;
        e = hipMemcpyAsync(wb.host + slot, wb.ctl + kWfCount0 + cur, sizeof(unsigned), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipEventRecord(wb.ev[slot], st);
        if (e != hipSuccess) return int(e);
        if (pending >= 0) {
            // The count of two read-backs ago; the batch launched since keeps the GPU busy meanwhile.
            e = hipEventSynchronize(wb.ev[pending]);
            if (e != hipSuccess) return int(e);
            known = wb.host[pending];
            if (known <= tail || known == 0u) break;
        }
        pending = slot;
        ++batch;
    }
    hipLaunchKernelGGL((k_wf_finish<S, E>), grid_for(finish_blocks), block, 0, st, sc, P, io, wb.state[cur], stride,
                       wb.ctl, cur, cnt);
    if (run) {
        run->iters = it;
        run->tail = known;
    }
    return int(hipGetLastError());
}

}  // namespace

size_t wave_slot_words(bool) { return size_t(Slot<PathLane>::kWords); }

int run_wave(int stack, bool shade, const DevScene& sc, const FrameParams& P, const GBufferSoA& gb,
             const ReservoirSoA& rm, float4* hdr, uint8_t* rgb, float* lum_var, const WaveBufs& wb,
             unsigned long long* cnt, hipStream_t st, WaveRun* run) {
    const PassIO io{gb, rm, hdr, rgb, lum_var};
    return with_stack(stack, [&](auto cap) {
        constexpr int S = decltype(cap)::value;
        return shade ? run_wavefront<S, PathEngine<S, true>>(sc, P, io, wb, cnt, st, run)
                     : run_wavefront<S, PathEngine<S, false>>(sc, P, io, wb, cnt, st, run);
    });
}

void launch_wave(int stack, bool shade, const DevScene& sc, const FrameParams& P, const GBufferSoA& gb,
                 const ReservoirSoA& rm, float4* hdr, uint8_t* rgb, float* lum_var, unsigned int* queue,
                 unsigned int* order, unsigned int* order_n, unsigned long long* cnt, hipStream_t st) {
    const PassIO io{gb, rm, hdr, rgb, lum_var};
    BandQueue q{queue, nullptr, nullptr};
    if (shade && order) {
        const int npx = P.width * P.height;
        hipLaunchKernelGGL(k_shade_order, dim3((npx + kBlock - 1) / kBlock), dim3(kBlock), 0, st, P, gb, hdr, lum_var,
                           order, order_n);
        q.order = order;
        q.order_n = order_n;
    }
    const bool deep = sc.nnodes > kDeepTreeNodes;
    with_stack(stack, [&](auto cap) {
        constexpr int S = decltype(cap)::value;
        if constexpr (S <= 48) {
            if (deep) {
                if (shade) launch_persistent<S, PathEngineDeep<S, true>>(sc, P, io, q, cnt, st);
                else launch_persistent<S, PathEngineDeep<S, false>>(sc, P, io, q, cnt, st);
                return;
            }
        }
        if (shade) launch_persistent<S, PathEngine<S, true>>(sc, P, io, q, cnt, st);
        else launch_persistent<S, PathEngine<S, false>>(sc, P, io, q, cnt, st);
    });
}

void launch_generate(int stack, const DevScene& sc, const FrameParams& P, const GBufferSoA& gb, const ReservoirSoA& ri,
                     unsigned int* queue, unsigned long long* cnt, hipStream_t st) {
    const PassIO io{gb, ri, nullptr, nullptr, nullptr};
    const BandQueue q{queue, nullptr, nullptr};
    with_stack(stack, [&](auto cap) {
        constexpr int S = decltype(cap)::value;
        launch_persistent<S, GenEngine<S>>(sc, P, io, q, cnt, st);
    });
}

}  // namespace rmh
