#pragma once
// Deterministic single-precision exponential for the a-trous edge-stopping
// weights.
//
// Host libm and the device math library round expf differently in the last
// place, so a denoised pixel computed on the host and on the device could
// differ in its low bits. Both sides therefore evaluate exp with this one
// routine, made of IEEE single-precision operations only (explicit fused
// multiply-adds, an exact scaling by a power of two), so host and device
// results are bit-identical:
//   * reduction x = k ln2 + r with k = rint(x / ln2) and ln2 split in two
//     (the high part 0x3F317200 has 16 significant bits, so k * ln2_hi is
//     exact for |k| < 2^8), |r| <= ln2/2 + tiny;
//   * the Taylor series of e^r to degree 7 (truncation below 6e-9 relative)
//     in Horner form;
//   * scaling by 2^k, in two exact steps where the result is subnormal so
//     that it is rounded once.
// The result is within 1 ulp of libm's exp for every float (exhaustive test).
// Special values: NaN propagates, x < -104 gives +0 (e^-104 is below half the
// smallest float subnormal), x > 89 gives +inf.
#if defined(__HIPCC__)
#define RM_EXP_HD __host__ __device__
#else
#define RM_EXP_HD
#endif

RM_EXP_HD inline float rm_expf(float x) {
    if (x != x) return x;
    if (x < -104.0f) return 0.0f;
    if (x > 89.0f) return __builtin_inff();
    const float ln2_hi = 0.693145751953125f;     // 0x3F317200
    const float ln2_lo = 1.428606765330187e-06f;  // ln2 - ln2_hi, to float precision
    const float kf = __builtin_rintf(x * 1.44269504088896341f);
    float r = __builtin_fmaf(-kf, ln2_hi, x);  // exact product, one rounding
    r = __builtin_fmaf(-kf, ln2_lo, r);
    float p = 1.98412698e-04f;               // 1/7!
    p = __builtin_fmaf(p, r, 1.38888889e-03f);  // 1/6!
    p = __builtin_fmaf(p, r, 8.33333333e-03f);  // 1/5!
    p = __builtin_fmaf(p, r, 4.16666667e-02f);  // 1/4!
    p = __builtin_fmaf(p, r, 1.66666667e-01f);  // 1/3!
    p = __builtin_fmaf(p, r, 0.5f);
    p = __builtin_fmaf(p, r, 1.0f);
    p = __builtin_fmaf(p, r, 1.0f);
    const int k = int(kf);
    if (k < -125) return __builtin_ldexpf(__builtin_ldexpf(p, k + 64), -64);
    return __builtin_ldexpf(p, k);
}
