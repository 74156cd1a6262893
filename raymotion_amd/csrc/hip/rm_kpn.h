// Kernel-predicting denoiser: network shape, parameter layout and the launchers
// of rm_kpn.hip. Not part of the reference (its denoiser is the a-trous filter).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "rm_layout.h"

namespace rmh {

// Per-pixel input features (NHWC, padded to 16 channels): log(1 + radiance)
// (3), albedo (3), shading normal (3), log(1 + depth) (1), log(1 + luminance
// variance) (1), screen motion in pixels / 8 (2), zero (3).
constexpr int kKpnFeat = 16;
constexpr int kKpnTaps = 25;  // 5x5 filter per pixel
constexpr int kKpnHidden = 64;
constexpr int kKpnLayers = 6;

struct KpnLayer {
    int cin, cout;  // padded channel counts (multiples of 16)
    int kp;         // padded reduction length: 9 * cin rounded up to 32
    size_t w_off;   // offset of the [cout][kp] weights in the parameter vector
    size_t b_off;   // offset of the cout biases
};

// Layers 16->64, 4 x 64->64 (ReLU), 64->32 (25 logits used, no ReLU).
struct KpnNet {
    KpnLayer layer[kKpnLayers];
    size_t nparam;
};
KpnNet kpn_net();

// conv3x3 (zero padding) as an implicit GEMM on v_mfma_f32_16x16x32_bf16.
// mode 0: relu(x * w + b) -> bf16; mode 1: x * w + b -> fp32; mode 2 (input
// gradient): (x * w) masked where act <= 0 -> bf16; mode 3 (the 64 -> 32
// head at inference): the logits x * w + b go straight into the softmax and
// 5x5 filter of `rad` (as launch_kpn_apply), out = float4 per pixel (w = 0).
void launch_conv(int cin, int cout, int mode, const __bf16* x, const __bf16* w, const float* b, const __bf16* act,
                 void* out, int n, int h, int wdt, hipStream_t st, const float* rad = nullptr);
// Weights of the input-gradient convolution: wd[ci][t * cout + co] = w[co][(8 - t) * cin + ci].
void launch_flip_weights(int cin, int cout, const __bf16* w, __bf16* wd, hipStream_t st);
// Weight and bias gradients of one layer into grad (at the layer's offsets):
// partial sums per pixel chunk, then a fixed-order reduction. `partial` holds
// at least kpn_wgrad_scratch(...) floats.
size_t kpn_wgrad_scratch(int cin, int cout, int pixels);
void launch_wgrad(int cin, int cout, const __bf16* x, const __bf16* dz, float* partial, float* grad_w, float* grad_b,
                  int n, int h, int wdt, hipStream_t st);
// Softmax over 25 logits, 5x5 filter of the noisy radiance (edge-clamped).
// Writes out ([P][3]) and out4 ([P], w = 0) where not null; with ref, also the
// loss mean((log(1 + out) - log(1 + ref))^2) scaled by inv_count * 3 pixels
// (added to loss[0]) and the logits' gradient dlogit [P][32] in bf16.
void launch_kpn_apply(const float* logits, const float* rad, const float* ref, float* out, float4* out4,
                      __bf16* dlogit, float* loss, float inv_count, int n, int h, int wdt, hipStream_t st);
void launch_adam(float* p, const float* g, float* m, float* v, __bf16* pb, size_t count, float lr, float beta1,
                 float beta2, float eps, int step, hipStream_t st);
void launch_to_bf16(const float* src, __bf16* dst, size_t count, hipStream_t st);
// Features and noisy radiance of a ReSTIR frame from its G-buffer and shading
// output: feat (fp32) and featb (bf16, the network's input) where not null.
void launch_features(const FrameParams& P, const GBufferSoA& gb, const float4* shade, const float* lum_var,
                     float* feat, float* rad, hipStream_t st, __bf16* featb = nullptr);
// Gathers batch images `idx` (each h * w pixels) from the dataset arrays.
void launch_gather(const __bf16* feat_all, const float* rad_all, const float* ref_all, const int32_t* idx, int batch,
                   int px_per_image, __bf16* feat, float* rad, float* ref, hipStream_t st);

}  // namespace rmh
