// Kernels of the kernel-predicting denoiser on gfx950: BF16 MFMA
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation) convolutions for the forward
// pass and both gradients, the per-pixel 5x5 kernel application with its
// loss and gradient, Adam, and the feature builder.
//
// Activations are NHWC in bf16: a pixel's 64 channels are one 128-byte line.
// A conv3x3 is a GEMM over (pixels) x (output channels) with K = 9 taps x cin;
// a lane's A fragment is 8 consecutive channels of one (shifted) pixel, one
// 16-byte load, and its B fragment 8 consecutive K of one output channel.
#include "rm_kpn.h"

#include <algorithm>

#ifndef RM_CONV16_BLOCKS
#define RM_CONV16_BLOCKS 3
#endif

namespace rmh {

KpnNet kpn_net() {
    KpnNet net{};
    const int cin[kKpnLayers] = {kKpnFeat, kKpnHidden, kKpnHidden, kKpnHidden, kKpnHidden, kKpnHidden};
    const int cout[kKpnLayers] = {kKpnHidden, kKpnHidden, kKpnHidden, kKpnHidden, kKpnHidden, 32};
    size_t off = 0;
    for (int l = 0; l < kKpnLayers; ++l) {
        KpnLayer& L = net.layer[l];
        L.cin = cin[l];
        L.cout = cout[l];
        L.kp = (9 * cin[l] + 31) / 32 * 32;
        L.w_off = off;
        off += size_t(L.cout) * size_t(L.kp);
        L.b_off = off;
        off += size_t(L.cout);
    }
    net.nparam = off;
    return net;
}

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__device__ __forceinline__ bf16x8 zero8() {
    bf16x8 z;
#pragma unroll
    for (int j = 0; j < 8; ++j) z[j] = (__bf16)0.0f;
    return z;
}

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Cross-lane shifts of A fragments within each row of 16 lanes (one lane per
// pixel of a 16-pixel tile): next() gives lane l the fragment of lane l + 1 of
// cur and lane 15 that of lane 0 of nxt (the pixel to the right); prev() lane
// l that of lane l - 1 and lane 0 that of lane 15 of prv (the pixel to the
// left). DPP row_shl:1 / row_shr:1, the row's end filled from a row_ror of
// the neighbouring tile.
typedef int i32x4 __attribute__((ext_vector_type(4)));

// One all-zero 64-channel bf16 pixel (128 bytes), never written: the load
// source of taps that fall outside the image.
__device__ __align__(16) uint32_t kZeroPixel[32];

// 16-byte load through a global-address-space pointer (a generic pointer
// would make it a flat load, which also counts against the LDS wait counter).
typedef __attribute__((address_space(1))) const bf16x8 global_bf16x8;
__device__ __forceinline__ bf16x8 load_global(const __bf16* p) {
    return *reinterpret_cast<global_bf16x8*>(reinterpret_cast<uintptr_t>(p));
}
constexpr int kDppRowShl1 = 0x101, kDppRowShr1 = 0x111, kDppRowRor1 = 0x121, kDppRowRor15 = 0x12F;

__device__ __forceinline__ bf16x8 frag_next(bf16x8 cur, bf16x8 nxt) {
    const i32x4 a = __builtin_bit_cast(i32x4, cur), b = __builtin_bit_cast(i32x4, nxt);
    i32x4 r;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int wrap = __builtin_amdgcn_mov_dpp(b[i], kDppRowRor15, 0xF, 0xF, false);
        r[i] = __builtin_amdgcn_update_dpp(wrap, a[i], kDppRowShl1, 0xF, 0xF, false);
    }
    return __builtin_bit_cast(bf16x8, r);
}

__device__ __forceinline__ bf16x8 frag_prev(bf16x8 cur, bf16x8 prv) {
    const i32x4 a = __builtin_bit_cast(i32x4, cur), b = __builtin_bit_cast(i32x4, prv);
    i32x4 r;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int wrap = __builtin_amdgcn_mov_dpp(b[i], kDppRowRor1, 0xF, 0xF, false);
        r[i] = __builtin_amdgcn_update_dpp(wrap, a[i], kDppRowShr1, 0xF, 0xF, false);
    }
    return __builtin_bit_cast(bf16x8, r);
}

// Softmax of a pixel's 25 logits lg[0..24] and the 5x5 filter of its noisy
// radiance: tap(t) points at the 3 radiance values of tap t (row t / 5 - 2,
// column t % 5 - 2, edge-clamped). w receives the normalised weights, o the
// filtered radiance. Shared by k_kpn_apply and the fused head convolution, so
// both give the same bits.
template <class LG, class TAP>
__device__ __forceinline__ void kpn_filter(LG lg, TAP tap, float w[kKpnTaps], float o[3]) {
    float mx = lg[0];
    for (int t = 1; t < kKpnTaps; ++t) mx = fmaxf(mx, lg[t]);
    float sum = 0.0f;
    for (int t = 0; t < kKpnTaps; ++t) {
        w[t] = __expf(lg[t] - mx);
        sum += w[t];
    }
    const float inv = 1.0f / sum;
    o[0] = o[1] = o[2] = 0.0f;
    for (int t = 0; t < kKpnTaps; ++t) {
        w[t] *= inv;
        const float* r = tap(t);
        o[0] += w[t] * r[0];
        o[1] += w[t] * r[1];
        o[2] += w[t] * r[2];
    }
}

// Wave-level ordering of LDS accesses between the lanes of one wave.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Stores a wave's 2-D output tile (pixel (x0 + q % 16, y0 + q / 16) for
// q = 0..63, COUT bf16 channels each) from the MFMA layout (value(j, t, r):
// pixel q = 16 j + 4 kq + r, channel 16 t + r16) as 16-byte vectors: one
// 16-channel group at a time goes through the wave's 2 KiB LDS buffer `tw`,
// then lane l writes the two 8-channel halves of pixels l / 2 and 32 + l / 2
// (half l % 2). Pixels outside the W x H image are not written.
template <int NT, class VAL>
__device__ __forceinline__ void store_tile_bf16(__bf16* tw, VAL value, __bf16* __restrict__ out, int x0, int y0,
                                                size_t img, int W, int H) {
    constexpr int COUT = 16 * NT;
    const int lane = threadIdx.x & 63, r16 = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) tw[(16 * j + 4 * kq + r) * 16 + r16] = value(j, t, r);
        wave_lds_sync();
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int q = 32 * c + (lane >> 1), hh = lane & 1;
            const int xx = x0 + (q & 15), yy = y0 + (q >> 4);
            const bf16x8 v = *reinterpret_cast<const bf16x8*>(tw + q * 16 + 8 * hh);
            if (xx < W && yy < H)
                *reinterpret_cast<bf16x8*>(out + ((img + size_t(yy)) * W + xx) * COUT + 16 * t + 8 * hh) = v;
        }
        wave_lds_sync();  // the next group overwrites tw only after every lane read it
    }
}

// Pixels per wave tile (kConvJ row tiles of 16), k-steps of A fragments
// loaded ahead (layers with fewer than 64 input channels; those with 64 keep
// one tap row ahead), and resident blocks per CU. Measured on the 1280x720
// network (ms for the 6 layers, all on the per-step path then): J = 4 with
// 1 / 2 / 3 steps ahead 1.20 / 1.18 / 1.15 (before the epilogue fix), J = 2
// 1.33; J = 8 spills.
constexpr int kConvJ = 4;
constexpr int kConvDepth = 2;
constexpr int kConvBlocksPerCU = 2;

// A block stages the layer's whole weight matrix in LDS once, already in the
// MFMA B-fragment order (k-step s, column tile t, lane: 16 bytes each, so a
// wave's fragment read is one contiguous 1 KiB line), then its 4 waves loop
// over tiles of 16 * kConvJ pixels x COUT channels. Per k-step a wave takes
// kConvJ A fragments (8 channels of one shifted pixel per lane, loaded from
// global memory through L1, or for 64 input channels shifted across lanes from
// the tap row's centre column) and reads COUT / 16 B fragments from LDS, feeding
// kConvJ * COUT / 16 MFMAs. Each output's sum runs over the k-steps in order,
// as a plain GEMM loop would. At most 72 KiB of LDS per block: 2 blocks per CU.
template <int CIN, int COUT, int MODE>
__global__ __launch_bounds__(256, kConvBlocksPerCU) void k_conv(const __bf16* __restrict__ X,
                                                              const __bf16* __restrict__ Wt,
                                                              const float* __restrict__ bias,
                                                              const __bf16* __restrict__ act,
                                                              void* __restrict__ out, int N, int H, int W) {
    constexpr int KP = (9 * CIN + 31) / 32 * 32;
    constexpr int KS = KP / 32;
    constexpr int NT = COUT / 16;
    constexpr int J = kConvJ;
    __shared__ bf16x8 wl[KS * NT * 64];
    for (int f = threadIdx.x; f < KS * NT * 64; f += 256) {
        const int l = f & 63, t = (f >> 6) % NT, s = (f >> 6) / NT;
        const int co = t * 16 + (l & 15), k = s * 32 + 8 * (l >> 4);
        wl[f] = k < 9 * CIN ? *reinterpret_cast<const bf16x8*>(Wt + size_t(co) * KP + k) : zero8();
    }
    __syncthreads();
    const int P = N * H * W;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r16 = lane & 15, kq = lane >> 4;
    const int tiles = (P + 16 * J - 1) / (16 * J);
    // The lane's output channels are 16 t + r16: their biases, read once.
    float bco[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) bco[t] = MODE == 2 ? 0.0f : bias[16 * t + r16];
    for (int tile = blockIdx.x * 4 + wave; tile < tiles; tile += gridDim.x * 4) {
        const int m0 = tile * 16 * J;
        // Per row tile: the lane's pixel and which of the 9 taps lie inside
        // the image (bit 3 (dy + 1) + (dx + 1)); none for pixels beyond P.
        const __bf16* base[J];
        unsigned taps[J];
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int p = m0 + 16 * j + r16;
            base[j] = X;
            taps[j] = 0;
            if (p < P) {
                const int rem = p % (H * W);
                const int y = rem / W, x = rem - (rem / W) * W;
                const unsigned xm = (x > 0 ? 1u : 0u) | 2u | (x < W - 1 ? 4u : 0u);
                taps[j] = (y > 0 ? xm : 0u) | (xm << 3) | (y < H - 1 ? xm << 6 : 0u);
                base[j] = X + size_t(p) * CIN;
            }
        }
        f32x4 acc[J][NT];
#pragma unroll
        for (int j = 0; j < J; ++j)
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[j][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        if constexpr (CIN == 64) {
            // k-step s = 2 tap + half. Per tap row dy the wave loads the centre
            // column (dx = 0, both channel halves) of its J tiles and the one
            // pixel on either side, and forms the dx = -1 / +1 fragments by
            // cross-lane shifts: each input line is fetched once per tap row
            // instead of three times. Steps still run in order s = 0 .. 17.
            // centre columns of tap row dyi (both halves), next row in flight
            // during this row's MFMAs
            auto load_rows = [&](int dyi, bf16x8 (&c)[2][J + 2]) {
                const long row = long(dyi - 1) * W * CIN;
#pragma unroll
                for (int jj = 0; jj < J + 2; ++jj) {
                    bool ok;
                    const __bf16* src;
                    if (jj >= 1 && jj <= J) {
                        ok = (taps[jj - 1] >> (3 * dyi + 1)) & 1u;
                        src = base[jj - 1];
                    } else {
                        const int q = m0 + 16 * (jj - 1) + r16;
                        ok = (jj == 0 ? r16 == 15 : r16 == 0) && q >= 0 && q < P;
                        if (ok) {
                            const int yq = (q % (H * W)) / W + dyi - 1;
                            ok = yq >= 0 && yq < H;
                        }
                        src = X + size_t(ok ? q : 0) * CIN;
                    }
#pragma unroll
                    for (int h = 0; h < 2; ++h)
                        c[h][jj] = ok ? *reinterpret_cast<const bf16x8*>(src + row + 32 * h + 8 * kq) : zero8();
                }
            };
            bf16x8 c[2][J + 2], cn[2][J + 2];  // [half][tile + 1]: tiles -1 and J hold one pixel each
            load_rows(0, c);
#pragma unroll 1
            for (int dyi = 0; dyi < 3; ++dyi) {
                if (dyi < 2) load_rows(dyi + 1, cn);
#pragma unroll
                for (int dxi = 0; dxi < 3; ++dxi)
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const int tap = 3 * dyi + dxi, s = 2 * tap + h;
                        bf16x8 a[J];
#pragma unroll
                        for (int j = 0; j < J; ++j) {
                            const bf16x8 v = dxi == 1 ? c[h][j + 1]
                                             : dxi == 2 ? frag_next(c[h][j + 1], c[h][j + 2])
                                                        : frag_prev(c[h][j + 1], c[h][j]);
                            a[j] = (taps[j] >> tap) & 1u ? v : zero8();
                        }
#pragma unroll
                        for (int t = 0; t < NT; ++t) {
                            const bf16x8 b = wl[(s * NT + t) * 64 + lane];
#pragma unroll
                            for (int j = 0; j < J; ++j) acc[j][t] = mfma(a[j], b, acc[j][t]);
                        }
                    }
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int jj = 0; jj < J + 2; ++jj) c[h][jj] = cn[h][jj];
            }
        } else {
        // A fragments of k-step s: 8 channels of the tap-shifted pixel, zero
        // outside the image and in the K padding (taps 9 and up).
        // Taps outside the image (and the K padding) load from kZeroPixel,
        // so no branch surrounds the loads (see k_conv64).
        auto load_a = [&](int s, bf16x8 (&a)[J]) {
            const int k = s * 32 + 8 * kq;
            const int tap = k / CIN, ci = k - tap * CIN;
            const long off = long((tap / 3 - 1) * W + (tap % 3 - 1)) * CIN + ci;
            const __bf16* zero = reinterpret_cast<const __bf16*>(kZeroPixel);
#pragma unroll
            for (int j = 0; j < J; ++j)
                a[j] = load_global((taps[j] >> tap) & 1u ? base[j] + off : zero);
        };
        // kConvDepth k-steps of A fragments in flight ahead of the MFMAs.
        bf16x8 a[kConvDepth + 1][J];
#pragma unroll
        for (int d = 0; d < kConvDepth; ++d)
            if (d < KS) load_a(d, a[d]);
#pragma unroll 1
        for (int s = 0; s < KS; ++s) {
            if (s + kConvDepth < KS) load_a(s + kConvDepth, a[kConvDepth]);
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const bf16x8 b = wl[(s * NT + t) * 64 + lane];
#pragma unroll
                for (int j = 0; j < J; ++j) acc[j][t] = mfma(a[0][j], b, acc[j][t]);
            }
#pragma unroll
            for (int d = 0; d < kConvDepth; ++d)
#pragma unroll
                for (int j = 0; j < J; ++j) a[d][j] = a[d + 1][j];
        }
        }
        // acc[j][t][r] is output pixel m0 + 16 j + 4 kq + r, channel 16 t + r16.
        const bool full = m0 + 16 * J <= P;
        if constexpr (MODE == 2) {
            // the activation mask of every output first (one bit each), then
            // the stores, so the loads are all in flight together
            static_assert(J * NT * 4 <= 64, "one mask bit per output of the lane");
            __bf16 av[J][NT][4];
#pragma unroll
            for (int j = 0; j < J; ++j)
#pragma unroll
                for (int t = 0; t < NT; ++t)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int m = min(m0 + 16 * j + 4 * kq + r, P - 1);  // beyond P: not stored
                        av[j][t][r] = act[size_t(m) * COUT + 16 * t + r16];
                    }
            unsigned long long on = 0;
#pragma unroll
            for (int j = 0; j < J; ++j)
#pragma unroll
                for (int t = 0; t < NT; ++t)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        on |= (unsigned long long)(float(av[j][t][r]) > 0.0f) << ((j * NT + t) * 4 + r);
#pragma unroll
            for (int j = 0; j < J; ++j)
#pragma unroll
                for (int t = 0; t < NT; ++t)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int m = m0 + 16 * j + 4 * kq + r;
                        if (!full && m >= P) continue;
                        static_cast<__bf16*>(out)[size_t(m) * COUT + 16 * t + r16] =
                            (on >> ((j * NT + t) * 4 + r)) & 1ull ? (__bf16)acc[j][t][r] : (__bf16)0.0f;
                    }
        } else {
#pragma unroll
            for (int j = 0; j < J; ++j)
#pragma unroll
                for (int t = 0; t < NT; ++t)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int m = m0 + 16 * j + 4 * kq + r;
                        if (!full && m >= P) continue;
                        const size_t o = size_t(m) * COUT + 16 * t + r16;
                        const float v = acc[j][t][r] + bco[t];
                        if (MODE == 0) static_cast<__bf16*>(out)[o] = (__bf16)fmaxf(v, 0.0f);
                        else static_cast<float*>(out)[o] = v;
                    }
        }
    }
}

// The 64-input-channel layers on 2-D tiles: a wave's tile is 4 image rows x
// 16 pixels (row tile j = image row 4 rg + j, lane r16 = pixel 16 seg + r16)
// x COUT channels, so the 3 tap rows of its 4 output rows are 6 image rows,
// each loaded once (both channel halves of the 16 centre pixels, and the
// pixel on either side for the row's left / right neighbours), instead of 3
// tap rows per 64-pixel strip. A loaded row is kept while the tap rows that
// use it run, its column shifts (dx = -1, +1) formed by DPP row shifts at
// each use: tap row dy of row tile j reads image row j + dy + 1 of the
// tile, so rows 0-3 serve dy = -1, rows 1-4 dy = 0 and rows 2-5 dy = +1
// (four rows kept, the next one loaded during the current tap row).
// Image rows and pixels outside the image load as zeros, which is the
// convolution's zero padding, so no per-tap masks are needed. Each output
// sums the k-steps s = 2 (3 dy + dx) + half in order, as in k_conv.
template <int COUT, int MODE>
__global__ __launch_bounds__(256, kConvBlocksPerCU) void k_conv64(const __bf16* __restrict__ X,
                                                                const __bf16* __restrict__ Wt,
                                                                const float* __restrict__ bias,
                                                                const __bf16* __restrict__ act,
                                                                void* __restrict__ out, int N, int H, int W,
                                                                const float* __restrict__ rad) {
    constexpr int CIN = 64;
    constexpr int KP = 9 * CIN;
    constexpr int KS = KP / 32;
    constexpr int NT = COUT / 16;
    static_assert(MODE != 3 || COUT == 32, "the fused filter takes the 32-channel head");
    __shared__ bf16x8 wl[KS * NT * 64];
    // MODE 3: each wave's tile of logits (64 pixels x 25), handed from the
    // MFMA layout (a pixel's channels across 16 lanes) to one lane per pixel,
    // and its radiance window, (16 + 4) x (4 + 4) pixels around the tile,
    // edge-clamped as the filter's taps are. Modes 0 and 2: each wave's 2 KiB
    // staging buffer of store_tile_bf16 (with the 72 KiB of 64 -> 64 weights,
    // two blocks take the 160 KiB of LDS exactly).
    constexpr int kWinW = 20, kWinH = 8;
    constexpr int kStageWords = MODE == 3 ? 4 * (64 * kKpnTaps + kWinW * kWinH * 3) : 4 * 64 * 16 / 2;
    __shared__ __align__(16) float stage[kStageWords];
    float* const lgs = stage;
    float* const rws = stage + 4 * 64 * kKpnTaps;
    __bf16* const tws = reinterpret_cast<__bf16*>(stage);
    for (int f = threadIdx.x; f < KS * NT * 64; f += 256) {
        const int l = f & 63, t = (f >> 6) % NT, s = (f >> 6) / NT;
        const int co = t * 16 + (l & 15), k = s * 32 + 8 * (l >> 4);
        wl[f] = *reinterpret_cast<const bf16x8*>(Wt + size_t(co) * KP + k);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r16 = lane & 15, kq = lane >> 4;
    const int segs = (W + 15) / 16, groups = (H + 3) / 4;
    const int tiles = N * groups * segs;
    float bco[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) bco[t] = MODE == 2 ? 0.0f : bias[16 * t + r16];
    // The 4 waves of a block take 4 neighbouring tiles of one row group.
    for (int tile = int(blockIdx.x) * 4 + wave; tile < tiles; tile += int(gridDim.x) * 4) {
        const int seg = tile % segs, rest = tile / segs;
        const int rg = rest % groups, n = rest / groups;
        const int x0 = seg * 16, y0 = rg * 4;
        const int x = x0 + r16;
        const size_t img = size_t(n) * H;
        // Fragments of image row yy: c[h], the lane's pixel for channel half
        // h; e[h], the row's ends in one register (lane 15: the pixel left of
        // x0, lane 0: the pixel right of x0 + 15), which is where frag_prev
        // and frag_next read them. Rows and pixels outside the image are
        // loaded from kZeroPixel: every lane loads, with no branch or select
        // after the load, so the compiler can wait for an older row while
        // this one is still in flight.
        auto load_row = [&](int yy, bf16x8 (&c)[2], bf16x8 (&e)[2]) {
            const bool row = yy >= 0 && yy < H;
            const __bf16* line = X + (img + size_t(row ? yy : 0)) * size_t(W) * CIN;
            const bool okc = row && x < W;
            const int xe = r16 == 15 ? x0 - 1 : x0 + 16;
            const bool oke = row && ((r16 == 15 && x0 > 0) || (r16 == 0 && x0 + 16 < W));
            const __bf16* zero = reinterpret_cast<const __bf16*>(kZeroPixel);
            const __bf16* pcen = (okc ? line + size_t(x) * CIN : zero) + 8 * kq;
            const __bf16* pend = (oke ? line + size_t(xe) * CIN : zero) + 8 * kq;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                c[h] = load_global(pcen + 32 * h);
                e[h] = load_global(pend + 32 * h);
            }
        };
        // rc[j], re[j]: tile row dyi + j (image row y0 - 1 + dyi + j), the
        // input of row tile j at tap row dyi
        bf16x8 rc[4][2], re[4][2];
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) load_row(y0 - 1 + rr, rc[rr], re[rr]);
        // MODE 3: the radiance window of the filter, (16 + 4) x (4 + 4) pixels
        // around the tile, edge-clamped, loaded now so that the loads complete
        // under the convolution (window position lane + 64 w, w = 0..2).
        float rwin[3][3];
        if constexpr (MODE == 3) {
#pragma unroll
            for (int w = 0; w < 3; ++w) {
                const int q = min(lane + 64 * w, kWinW * kWinH - 1);
                const int xx = min(max(x0 - 2 + q % kWinW, 0), W - 1);
                const int yy = min(max(y0 - 2 + q / kWinW, 0), H - 1);
                const float* src = rad + ((img + size_t(yy)) * W + xx) * 3;
                rwin[w][0] = src[0];
                rwin[w][1] = src[1];
                rwin[w][2] = src[2];
            }
        }
        {
            f32x4 acc[4][NT];
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int t = 0; t < NT; ++t) acc[j][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
            // Tap row dyi of row tile j reads tile row dyi + j: rows 0-3 (rc,
            // re) and rows 4, 5, each loaded at the start of the tap row
            // before its first use, so it is in flight during that tap row.
            // The scheduling barriers keep each tap row's instructions
            // together (unrolled freely, the compiler hoists the later rows'
            // LDS reads and runs out of registers).
            bf16x8 r4c[2], r4e[2], r5c[2], r5e[2];
#pragma unroll
            for (int dyi = 0; dyi < 3; ++dyi) {
                if (dyi == 0) load_row(y0 + 3, r4c, r4e);
                if (dyi == 1) load_row(y0 + 4, r5c, r5e);
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    const int dxi = q >> 1, h = q & 1, s = 6 * dyi + q;
                    bf16x8 a[4];  // the column shift dx = dxi - 1 of each row tile's input row
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int i = dyi + j;
                        const bf16x8 c = i < 4 ? rc[i][h] : i == 4 ? r4c[h] : r5c[h];
                        const bf16x8 e = i < 4 ? re[i][h] : i == 4 ? r4e[h] : r5e[h];
                        a[j] = dxi == 1 ? c : dxi == 0 ? frag_prev(c, e) : frag_next(c, e);
                    }
#pragma unroll
                    for (int t = 0; t < NT; ++t) {
                        const bf16x8 b = wl[(s * NT + t) * 64 + lane];
#pragma unroll
                        for (int j = 0; j < 4; ++j) acc[j][t] = mfma(a[j], b, acc[j][t]);
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
            }
            // acc[j][t][r] is pixel (x0 + 4 kq + r, y0 + j), channel 16 t + r16.
            auto pixel_of = [&](int j, int r) -> long {
                const int xx = x0 + 4 * kq + r, yy = y0 + j;
                return (xx < W && yy < H) ? long(img + size_t(yy)) * W + xx : -1;
            };
            if constexpr (MODE == 3) {
                // Logits (bias added as in mode 1) into this wave's LDS tile,
                // then lane q filters pixel (x0 + q % 16, y0 + q / 16) exactly
                // as k_kpn_apply does from the fp32 logits.
                float* lw = lgs + wave * 64 * kKpnTaps;
                float* rw = rws + wave * kWinW * kWinH * 3;
#pragma unroll
                for (int w = 0; w < 3; ++w) {
                    const int q = lane + 64 * w;
                    if (q < kWinW * kWinH) {
                        rw[3 * q] = rwin[w][0];
                        rw[3 * q + 1] = rwin[w][1];
                        rw[3 * q + 2] = rwin[w][2];
                    }
                }
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int t = 0; t < NT; ++t)
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int c = 16 * t + r16;
                            if (c < kKpnTaps) lw[(16 * j + 4 * kq + r) * kKpnTaps + c] = acc[j][t][r] + bco[t];
                        }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                const int xq = x0 + (lane & 15), yq = y0 + (lane >> 4);
                if (xq < W && yq < H) {
                    float wt[kKpnTaps], o[3];
                    const int lx = lane & 15, ly = lane >> 4;
                    kpn_filter(lw + lane * kKpnTaps,
                               [&](int t) { return rw + ((ly + t / 5) * kWinW + lx + t % 5) * 3; }, wt, o);
                    static_cast<float4*>(out)[size_t(img + size_t(yq)) * W + xq] = make_float4(o[0], o[1], o[2], 0.0f);
                }
                // the next tile's logits overwrite this tile's only after every lane read them
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            } else if constexpr (MODE == 2) {
                static_assert(4 * NT * 4 <= 64, "one mask bit per output of the lane");
                __bf16 av[4][NT][4];
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int t = 0; t < NT; ++t)
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const long m = pixel_of(j, r);
                            av[j][t][r] = act[size_t(m < 0 ? 0 : m) * COUT + 16 * t + r16];
                        }
                unsigned long long on = 0;
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int t = 0; t < NT; ++t)
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            on |= (unsigned long long)(float(av[j][t][r]) > 0.0f) << ((j * NT + t) * 4 + r);
                store_tile_bf16<NT>(
                    tws + wave * 64 * 16,
                    [&](int j, int t, int r) {
                        return (on >> ((j * NT + t) * 4 + r)) & 1ull ? (__bf16)acc[j][t][r] : (__bf16)0.0f;
                    },
                    static_cast<__bf16*>(out), x0, y0, img, W, H);
            } else if constexpr (MODE == 0) {
                store_tile_bf16<NT>(
                    tws + wave * 64 * 16,
                    [&](int j, int t, int r) { return (__bf16)fmaxf(acc[j][t][r] + bco[t], 0.0f); },
                    static_cast<__bf16*>(out), x0, y0, img, W, H);
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int t = 0; t < NT; ++t)
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const long m = pixel_of(j, r);
                            if (m < 0) continue;
                            static_cast<float*>(out)[size_t(m) * COUT + 16 * t + r16] = acc[j][t][r] + bco[t];
                        }
            }
        }
    }
}

// The 16-input-channel first layer on the same 2-D tiles (4 image rows x 16
// pixels per wave). K = 9 taps x 16 channels, padded to 5 k-steps of 32: k-step
// s holds tap 2s in lanes kq = 0, 1 (channels 0-7, 8-15) and tap 2s + 1 in
// lanes kq = 2, 3, as in k_conv, so every output sums the same MFMAs in the
// same order and the results equal k_conv's bit for bit. Each of the 6 image
// rows a tile needs is loaded once per lane (the lane's 8 channels of its
// pixel, and of the row's end pixels), where k_conv loads every tap's pixel:
// a k-step's fragment for row tile j is the input row j + dy + 1 shifted by
// dx across lanes, formed for both of the k-step's taps and selected by kq.
template <int COUT, int MODE>
__global__ __launch_bounds__(256, RM_CONV16_BLOCKS) void k_conv16(const __bf16* __restrict__ X,
                                                                const __bf16* __restrict__ Wt,
                                                                const float* __restrict__ bias,
                                                                void* __restrict__ out, int N, int H, int W) {
    static_assert(MODE == 0, "forward (bias + ReLU) only");
    constexpr int CIN = 16;
    constexpr int KP = (9 * CIN + 31) / 32 * 32;
    constexpr int KS = KP / 32;
    constexpr int NT = COUT / 16;
    __shared__ bf16x8 wl[KS * NT * 64];
    __shared__ __align__(16) __bf16 tws[4 * 64 * 16];  // store_tile_bf16 staging, per wave
    for (int f = threadIdx.x; f < KS * NT * 64; f += 256) {
        const int l = f & 63, t = (f >> 6) % NT, s = (f >> 6) / NT;
        const int co = t * 16 + (l & 15), k = s * 32 + 8 * (l >> 4);
        wl[f] = k < 9 * CIN ? *reinterpret_cast<const bf16x8*>(Wt + size_t(co) * KP + k) : zero8();
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r16 = lane & 15, kq = lane >> 4;
    const int half = kq & 1, upper = kq >> 1;  // channel half, and which of the k-step's two taps
    const int segs = (W + 15) / 16, groups = (H + 3) / 4;
    const int tiles = N * groups * segs;
    float bco[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) bco[t] = bias[16 * t + r16];
    for (int tile = int(blockIdx.x) * 4 + wave; tile < tiles; tile += int(gridDim.x) * 4) {
        const int seg = tile % segs, rest = tile / segs;
        const int rg = rest % groups, n = rest / groups;
        const int x0 = seg * 16, y0 = rg * 4;
        const int x = x0 + r16;
        const size_t img = size_t(n) * H;
        // rc[i]: the lane's channel half of pixel x in image row y0 - 1 + i;
        // re[i]: the row's end pixels (lane 15: x0 - 1, lane 0: x0 + 16).
        bf16x8 rc[6], re[6];
        const __bf16* zero = reinterpret_cast<const __bf16*>(kZeroPixel);
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            const int yy = y0 - 1 + i;
            const bool row = yy >= 0 && yy < H;
            const __bf16* line = X + (img + size_t(row ? yy : 0)) * size_t(W) * CIN;
            const bool okc = row && x < W;
            const int xe = r16 == 15 ? x0 - 1 : x0 + 16;
            const bool oke = row && ((r16 == 15 && x0 > 0) || (r16 == 0 && x0 + 16 < W));
            rc[i] = load_global((okc ? line + size_t(x) * CIN : zero) + 8 * half);
            re[i] = load_global((oke ? line + size_t(xe) * CIN : zero) + 8 * half);
        }
        f32x4 acc[4][NT];
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[j][t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const int ta = 2 * s, tb = 2 * s + 1;  // the k-step's taps (tb = 9 is padding)
            bf16x8 a[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                auto shifted = [&](int tap) -> bf16x8 {
                    const int dyi = tap / 3, dxi = tap % 3;
                    const bf16x8 c = rc[j + dyi], e = re[j + dyi];
                    return dxi == 1 ? c : dxi == 0 ? frag_prev(c, e) : frag_next(c, e);
                };
                const bf16x8 fa = shifted(ta);
                const bf16x8 fb = tb < 9 ? shifted(tb) : zero8();
                a[j] = upper ? fb : fa;
            }
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const bf16x8 b = wl[(s * NT + t) * 64 + lane];
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[j][t] = mfma(a[j], b, acc[j][t]);
            }
        }
        // acc[j][t][r] is pixel (x0 + 4 kq + r, y0 + j), channel 16 t + r16.
        store_tile_bf16<NT>(
            tws + wave * 64 * 16, [&](int j, int t, int r) { return (__bf16)fmaxf(acc[j][t][r] + bco[t], 0.0f); },
            static_cast<__bf16*>(out), x0, y0, img, W, H);
    }
}

template <int CIN, int COUT>
__global__ void k_flip(const __bf16* __restrict__ w, __bf16* __restrict__ wd) {
    constexpr int KP = (9 * CIN + 31) / 32 * 32;
    constexpr int KPD = (9 * COUT + 31) / 32 * 32;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;  // over CIN * KPD
    if (i >= CIN * KPD) return;
    const int ci = i / KPD, k = i - ci * KPD;
    __bf16 v = (__bf16)0.0f;
    if (k < 9 * COUT) {
        const int t = k / COUT, co = k - t * COUT;
        v = w[size_t(co) * KP + (8 - t) * CIN + ci];
    }
    wd[i] = v;
}

constexpr int kWgradChunk = 2048;  // pixels per wgrad block

// Weight gradient of one tap: C[ci][co] = sum_p x[p + off(tap)][ci] dz[p][co]
// over a chunk of pixels; 32-pixel K steps staged transposed in LDS.
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_wgrad(const __bf16* __restrict__ X, const __bf16* __restrict__ dZ,
                                               float* __restrict__ partial, int N, int H, int W) {
    constexpr int MT = CIN / 16, NT = COUT / 16, TILES = MT * NT, TPW = (TILES + 3) / 4;
    constexpr int PITCH = 40;  // 32 pixels + 8 pad (keeps 16-byte alignment)
    __shared__ __align__(16) __bf16 sx[CIN * PITCH];
    __shared__ __align__(16) __bf16 sd[COUT * PITCH];
    const int tap = blockIdx.y, dy = tap / 3 - 1, dx = tap % 3 - 1;
    const int P = N * H * W;
    const int p_begin = blockIdx.x * kWgradChunk, p_end = min(P, p_begin + kWgradChunk);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, r16 = lane & 15, kq = lane >> 4;
    f32x4 acc[TPW];
#pragma unroll
    for (int i = 0; i < TPW; ++i) acc[i] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int p0 = p_begin; p0 < p_end; p0 += 32) {
        for (int v = threadIdx.x; v < 32 * CIN / 8; v += 256) {
            const int pl = v / (CIN / 8), cg = (v - pl * (CIN / 8)) * 8;
            const int p = p0 + pl;
            bf16x8 val = zero8();
            if (p < p_end) {
                const int n = p / (H * W), rem = p - n * H * W, y = rem / W, x = rem - y * W;
                const int yy = y + dy, xx = x + dx;
                if (yy >= 0 && yy < H && xx >= 0 && xx < W)
                    val = *reinterpret_cast<const bf16x8*>(X + ((size_t(n) * H + yy) * W + xx) * CIN + cg);
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) sx[(cg + j) * PITCH + pl] = val[j];
        }
        for (int v = threadIdx.x; v < 32 * COUT / 8; v += 256) {
            const int pl = v / (COUT / 8), cg = (v - pl * (COUT / 8)) * 8;
            const int p = p0 + pl;
            bf16x8 val = zero8();
            if (p < p_end) val = *reinterpret_cast<const bf16x8*>(dZ + size_t(p) * COUT + cg);
#pragma unroll
            for (int j = 0; j < 8; ++j) sd[(cg + j) * PITCH + pl] = val[j];
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < TPW; ++i) {
            const int tile = wave + 4 * i;
            if (tile < TILES) {
                const int mt = tile / NT, nt = tile - mt * NT;
                const bf16x8 a = *reinterpret_cast<const bf16x8*>(&sx[(mt * 16 + r16) * PITCH + 8 * kq]);
                const bf16x8 b = *reinterpret_cast<const bf16x8*>(&sd[(nt * 16 + r16) * PITCH + 8 * kq]);
                acc[i] = mfma(a, b, acc[i]);
            }
        }
        __syncthreads();
    }
    // partial[tap][chunk][co][ci]; acc[i][r] is ci = 16 mt + 4 kq + r, co = 16 nt + r16.
    float* base = partial + (size_t(tap) * gridDim.x + blockIdx.x) * (CIN * COUT);
#pragma unroll
    for (int i = 0; i < TPW; ++i) {
        const int tile = wave + 4 * i;
        if (tile >= TILES) continue;
        const int mt = tile / NT, nt = tile - mt * NT;
        const int co = nt * 16 + r16;
#pragma unroll
        for (int r = 0; r < 4; ++r) base[co * CIN + mt * 16 + 4 * kq + r] = acc[i][r];
    }
}

// grad_w[co][tap * CIN + ci] = sum over chunks of partial[tap][chunk][co][ci].
template <int CIN, int COUT>
__global__ void k_wgrad_reduce(const float* __restrict__ partial, float* __restrict__ grad_w, int chunks) {
    constexpr int KP = (9 * CIN + 31) / 32 * 32;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;  // over 9 * COUT * CIN
    if (i >= 9 * COUT * CIN) return;
    const int tap = i / (COUT * CIN), rem = i - tap * COUT * CIN, co = rem / CIN, ci = rem - co * CIN;
    const float* src = partial + size_t(tap) * chunks * (CIN * COUT) + co * CIN + ci;
    float s = 0.0f;
    for (int c = 0; c < chunks; ++c) s += src[size_t(c) * (CIN * COUT)];
    grad_w[size_t(co) * KP + tap * CIN + ci] = s;
}

// Bias gradient partials: partial[chunk][co] = sum of dz over the chunk's pixels.
template <int COUT>
__global__ __launch_bounds__(256) void k_bgrad(const __bf16* __restrict__ dZ, float* __restrict__ partial, int P) {
    constexpr int ROWS = 256 / COUT;
    __shared__ float s[256];
    const int c = threadIdx.x % COUT, row = threadIdx.x / COUT;
    const int p_begin = blockIdx.x * kWgradChunk, p_end = min(P, p_begin + kWgradChunk);
    float acc = 0.0f;
    for (int p = p_begin + row; p < p_end; p += ROWS) acc += float(dZ[size_t(p) * COUT + c]);
    s[threadIdx.x] = acc;
    __syncthreads();
    if (row == 0) {
        for (int r = 1; r < ROWS; ++r) acc += s[r * COUT + c];
        partial[size_t(blockIdx.x) * COUT + c] = acc;
    }
}

template <int COUT>
__global__ void k_bgrad_reduce(const float* __restrict__ partial, float* __restrict__ grad_b, int chunks) {
    const int c = threadIdx.x;
    if (c >= COUT) return;
    float s = 0.0f;
    for (int k = 0; k < chunks; ++k) s += partial[size_t(k) * COUT + c];
    grad_b[c] = s;
}

__global__ __launch_bounds__(256) void k_kpn_apply(const float* __restrict__ logits, const float* __restrict__ rad,
                                                   const float* __restrict__ ref, float* __restrict__ out,
                                                   float4* __restrict__ out4, __bf16* __restrict__ dlogit,
                                                   float* __restrict__ loss, float inv_count, int N, int H, int W) {
    __shared__ float red[256];
    const int P = N * H * W;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    float my_loss = 0.0f;
    if (p < P) {
        const int n = p / (H * W), rem = p - n * H * W, y = rem / W, x = rem - y * W;
        float w[kKpnTaps], o[3];
        kpn_filter(logits + size_t(p) * 32,
                   [&](int t) {
                       const int yy = min(max(y + t / 5 - 2, 0), H - 1), xx = min(max(x + t % 5 - 2, 0), W - 1);
                       return rad + ((size_t(n) * H + yy) * W + xx) * 3;
                   },
                   w, o);
        if (out) {
            out[size_t(p) * 3] = o[0];
            out[size_t(p) * 3 + 1] = o[1];
            out[size_t(p) * 3 + 2] = o[2];
        }
        if (out4) out4[p] = make_float4(o[0], o[1], o[2], 0.0f);
        if (ref) {
            float g[3];
            for (int c = 0; c < 3; ++c) {
                const float d = log1pf(fmaxf(o[c], 0.0f)) - log1pf(fmaxf(ref[size_t(p) * 3 + c], 0.0f));
                my_loss += d * d;
                g[c] = 2.0f * d / (1.0f + fmaxf(o[c], 0.0f)) * inv_count;
            }
            float dw[kKpnTaps], dot = 0.0f;
            for (int t = 0; t < kKpnTaps; ++t) {
                const int yy = min(max(y + t / 5 - 2, 0), H - 1), xx = min(max(x + t % 5 - 2, 0), W - 1);
                const float* r = rad + ((size_t(n) * H + yy) * W + xx) * 3;
                dw[t] = g[0] * r[0] + g[1] * r[1] + g[2] * r[2];
                dot += w[t] * dw[t];
            }
            __bf16* dl = dlogit + size_t(p) * 32;
            for (int t = 0; t < kKpnTaps; ++t) dl[t] = (__bf16)(w[t] * (dw[t] - dot));
            for (int t = kKpnTaps; t < 32; ++t) dl[t] = (__bf16)0.0f;
        }
    }
    if (ref && loss) {
        red[threadIdx.x] = my_loss;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) atomicAdd(loss, red[0] * inv_count);
    }
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       __bf16* __restrict__ pb, size_t count, float lr, float b1, float b2, float eps, float c1,
                       float c2) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    p[i] = pi;
    pb[i] = (__bf16)pi;
}

__global__ void k_to_bf16(const float* __restrict__ s, __bf16* __restrict__ d, size_t count) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) d[i] = (__bf16)s[i];
}

__global__ void k_features(FrameParams P, GBufferSoA gb, const float4* __restrict__ shade,
                           const float* __restrict__ lum_var, float* __restrict__ feat, float* __restrict__ rad,
                           __bf16* __restrict__ featb) {
    const int px = blockIdx.x * blockDim.x + threadIdx.x;
    if (px >= P.width * P.height) return;
    const float4 g0 = gb.p0[px], g1 = gb.p1[px], g2 = gb.p2[px];
    const int kind = __float_as_int(g0.w);
    const float4 c = shade[px];
    float f[kKpnFeat];
    for (int i = 0; i < kKpnFeat; ++i) f[i] = 0.0f;
    f[0] = log1pf(fmaxf(c.x, 0.0f));
    f[1] = log1pf(fmaxf(c.y, 0.0f));
    f[2] = log1pf(fmaxf(c.z, 0.0f));
    if (kind != 0) {
        f[3] = g2.x; f[4] = g2.y; f[5] = g2.z;
        f[6] = g1.x; f[7] = g1.y; f[8] = g1.z;
        const float dx = g0.x - P.cam.x, dy = g0.y - P.cam.y, dz = g0.z - P.cam.z;
        f[9] = log1pf(sqrtf(dx * dx + dy * dy + dz * dz));
        // Screen motion: where the hit projected through the previous camera.
        if (P.has_history) {
            const rm_camera& q = P.prev;
            const float qx = g0.x - q.x, qy = g0.y - q.y, qz = g0.z - q.z;
            const float cx = q.m0 * qx + q.m3 * qy + q.m6 * qz, cy = q.m1 * qx + q.m4 * qy + q.m7 * qz,
                        cz = q.m2 * qx + q.m5 * qy + q.m8 * qz;
            if (cz > 1e-6f) {
                const float sx = q.focal * cx / cz + float(P.width) * 0.5f;
                const float sy = float(P.height) * 0.5f - q.focal * cy / cz;
                const float x = float(px % P.width) + 0.5f, y = float(px / P.width) + 0.5f;
                f[11] = fminf(fmaxf((sx - x) * 0.125f, -8.0f), 8.0f);
                f[12] = fminf(fmaxf((sy - y) * 0.125f, -8.0f), 8.0f);
            }
        }
    }
    f[10] = lum_var ? log1pf(fmaxf(lum_var[px], 0.0f)) : 0.0f;
    if (feat)
        for (int i = 0; i < kKpnFeat; ++i) feat[size_t(px) * kKpnFeat + i] = f[i];
    if (featb) {  // the network's input, converted as k_to_bf16 converts
        bf16x8 lo, hi;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            lo[i] = (__bf16)f[i];
            hi[i] = (__bf16)f[8 + i];
        }
        bf16x8* d = reinterpret_cast<bf16x8*>(featb + size_t(px) * kKpnFeat);
        d[0] = lo;
        d[1] = hi;
    }
    rad[size_t(px) * 3] = fmaxf(c.x, 0.0f);
    rad[size_t(px) * 3 + 1] = fmaxf(c.y, 0.0f);
    rad[size_t(px) * 3 + 2] = fmaxf(c.z, 0.0f);
}

__global__ void k_gather(const __bf16* __restrict__ feat_all, const float* __restrict__ rad_all,
                         const float* __restrict__ ref_all, const int32_t* __restrict__ idx, int batch, int ppi,
                         __bf16* __restrict__ feat, float* __restrict__ rad, float* __restrict__ ref) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;  // over batch * ppi pixels
    if (i >= size_t(batch) * ppi) return;
    const int b = int(i / ppi);
    const size_t src = size_t(idx[b]) * ppi + (i - size_t(b) * ppi);
    const bf16x8* fs = reinterpret_cast<const bf16x8*>(feat_all + src * kKpnFeat);
    bf16x8* fd = reinterpret_cast<bf16x8*>(feat + i * kKpnFeat);
    fd[0] = fs[0];
    fd[1] = fs[1];
    for (int c = 0; c < 3; ++c) {
        rad[i * 3 + c] = rad_all[src * 3 + c];
        ref[i * 3 + c] = ref_all[src * 3 + c];
    }
}

unsigned blocks_of(size_t n, unsigned b) { return unsigned((n + b - 1) / b); }

}  // namespace

void launch_conv(int cin, int cout, int mode, const __bf16* x, const __bf16* w, const float* b, const __bf16* act,
                 void* out, int n, int h, int wdt, hipStream_t st, const float* rad) {
    const size_t P = size_t(n) * h * wdt;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    // 64 input channels, and the 16-channel forward layer: 2-D tiles of 4
    // rows x 16 pixels (k_conv64, k_conv16); else strips of 16 * kConvJ
    // pixels (k_conv).
    const bool tiles2d = cin == 64 || (cin == 16 && mode == 0);
    const size_t tiles = tiles2d ? size_t(n) * size_t((h + 3) / 4) * size_t((wdt + 15) / 16)
                                 : (P + 16 * kConvJ - 1) / (16 * kConvJ);
    // The first layer (161 VGPRs, 28 KiB of LDS) fits 3 blocks per CU; more
    // waves keep more of its rows in flight (it is bound by HBM traffic).
    const int per_cu = cin == 16 && mode == 0 ? RM_CONV16_BLOCKS : kConvBlocksPerCU;
    const dim3 grid(unsigned(std::min<size_t>(blocks_of(tiles, 4), size_t(cus) * per_cu))), block(256);
#define RM_CONV(CI, CO, MO) \
    hipLaunchKernelGGL((k_conv<CI, CO, MO>), grid, block, 0, st, x, w, b, act, out, n, h, wdt)
#define RM_CONV64(CO, MO) \
    hipLaunchKernelGGL((k_conv64<CO, MO>), grid, block, 0, st, x, w, b, act, out, n, h, wdt, rad)
    if (cin == 16 && cout == 64 && mode == 0)
        hipLaunchKernelGGL((k_conv16<64, 0>), grid, block, 0, st, x, w, b, out, n, h, wdt);
    else if (cin == 64 && cout == 64 && mode == 0) RM_CONV64(64, 0);
    else if (cin == 64 && cout == 32 && mode == 1) RM_CONV64(32, 1);
    else if (cin == 64 && cout == 32 && mode == 3) RM_CONV64(32, 3);
    else if (cin == 64 && cout == 64 && mode == 2) RM_CONV64(64, 2);
    else if (cin == 32 && cout == 64 && mode == 2) RM_CONV(32, 64, 2);
#undef RM_CONV
#undef RM_CONV64
}

void launch_flip_weights(int cin, int cout, const __bf16* w, __bf16* wd, hipStream_t st) {
    const int kpd = (9 * cout + 31) / 32 * 32;
    const dim3 grid(blocks_of(size_t(cin) * kpd, 256)), block(256);
    if (cin == 64 && cout == 64) hipLaunchKernelGGL((k_flip<64, 64>), grid, block, 0, st, w, wd);
    else if (cin == 64 && cout == 32) hipLaunchKernelGGL((k_flip<64, 32>), grid, block, 0, st, w, wd);
}

size_t kpn_wgrad_scratch(int cin, int cout, int pixels) {
    const size_t chunks = (size_t(pixels) + kWgradChunk - 1) / kWgradChunk;
    return std::max(9 * chunks * size_t(cin) * cout, chunks * size_t(cout));
}

void launch_wgrad(int cin, int cout, const __bf16* x, const __bf16* dz, float* partial, float* grad_w, float* grad_b,
                  int n, int h, int wdt, hipStream_t st) {
    const int P = n * h * wdt;
    const int chunks = (P + kWgradChunk - 1) / kWgradChunk;
    const dim3 grid(chunks, 9), block(256);
    const dim3 rgrid(blocks_of(size_t(9) * cin * cout, 256));
#define RM_WGRAD(CI, CO)                                                                               \
    do {                                                                                               \
        hipLaunchKernelGGL((k_wgrad<CI, CO>), grid, block, 0, st, x, dz, partial, n, h, wdt);          \
        hipLaunchKernelGGL((k_wgrad_reduce<CI, CO>), rgrid, dim3(256), 0, st, partial, grad_w, chunks); \
        hipLaunchKernelGGL((k_bgrad<CO>), dim3(chunks), block, 0, st, dz, partial, P);                 \
        hipLaunchKernelGGL((k_bgrad_reduce<CO>), dim3(1), dim3(64), 0, st, partial, grad_b, chunks);   \
    } while (0)
    if (cin == 16 && cout == 64) RM_WGRAD(16, 64);
    else if (cin == 64 && cout == 64) RM_WGRAD(64, 64);
    else if (cin == 64 && cout == 32) RM_WGRAD(64, 32);
#undef RM_WGRAD
}

void launch_kpn_apply(const float* logits, const float* rad, const float* ref, float* out, float4* out4,
                      __bf16* dlogit, float* loss, float inv_count, int n, int h, int wdt, hipStream_t st) {
    const size_t P = size_t(n) * h * wdt;
    hipLaunchKernelGGL(k_kpn_apply, dim3(blocks_of(P, 256)), dim3(256), 0, st, logits, rad, ref, out, out4, dlogit,
                       loss, inv_count, n, h, wdt);
}

void launch_adam(float* p, const float* g, float* m, float* v, __bf16* pb, size_t count, float lr, float beta1,
                 float beta2, float eps, int step, hipStream_t st) {
    const float c1 = 1.0f / (1.0f - powf(beta1, float(step)));
    const float c2 = 1.0f / (1.0f - powf(beta2, float(step)));
    hipLaunchKernelGGL(k_adam, dim3(blocks_of(count, 256)), dim3(256), 0, st, p, g, m, v, pb, count, lr, beta1, beta2,
                       eps, c1, c2);
}

void launch_to_bf16(const float* src, __bf16* dst, size_t count, hipStream_t st) {
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks_of(count, 256)), dim3(256), 0, st, src, dst, count);
}

void launch_features(const FrameParams& P, const GBufferSoA& gb, const float4* shade, const float* lum_var,
                     float* feat, float* rad, hipStream_t st, __bf16* featb) {
    const size_t npx = size_t(P.width) * P.height;
    hipLaunchKernelGGL(k_features, dim3(blocks_of(npx, 256)), dim3(256), 0, st, P, gb, shade, lum_var, feat, rad,
                       featb);
}

void launch_gather(const __bf16* feat_all, const float* rad_all, const float* ref_all, const int32_t* idx, int batch,
                   int px_per_image, __bf16* feat, float* rad, float* ref, hipStream_t st) {
    const size_t n = size_t(batch) * px_per_image;
    hipLaunchKernelGGL(k_gather, dim3(blocks_of(n, 256)), dim3(256), 0, st, feat_all, rad_all, ref_all, idx, batch,
                       px_per_image, feat, rad, ref);
}

}  // namespace rmh
