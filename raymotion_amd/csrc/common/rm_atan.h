#pragma once
// Deterministic single-precision arctangent for the hue shader's azimuth.
//
// Host libm and the device math library round atanf differently in the last
// place, which moves a hue across a colour band edge now and then. Both the
// HIP kernels and the host code therefore evaluate atan with this one routine,
// made of IEEE single-precision operations only (explicit fused multiply-adds,
// IEEE division), so host and device results are bit-identical:
//   * a = |x|; for a > 1, atan(a) = pi/2 - atan(1/a);
//   * for a > 2 - sqrt(3) (tan 15 degrees), the addition formula
//     atan(a) = pi/6 + atan((a - c) / (1 + a c)) with c = 1/sqrt(3) leaves an
//     argument |t| <= 2 - sqrt(3);
//   * the Taylor series t - t^3/3 + t^5/5 - ... to degree 15 (truncation
//     below 5e-10 relative for |t| <= 0.268) in Horner form;
//   * pi/6 and pi/2 added as two-term constants (high part last).
// Single precision keeps the register footprint small in the kernels that
// inline it. The result is within 2 ulp of libm's atan for every float
// (exhaustive test). NaN propagates; atan(+-inf) = +-pi/2 rounded to float.
#if defined(__HIPCC__)
#define RM_ATAN_HD __host__ __device__
#else
#define RM_ATAN_HD
#endif

RM_ATAN_HD inline float rm_atanf(float x) {
    if (x != x) return x;
    const float pi_2_hi = 1.57079637e+00f, pi_2_lo = -4.37113883e-08f;  // pi/2 = hi + lo
    const float pi_6_hi = 5.23598790e-01f, pi_6_lo = -1.45704622e-08f;  // pi/6 = hi + lo
    const float inv_sqrt3 = 0.577350269f;
    const float tan15 = 0.267949194f;  // 2 - sqrt(3)
    float a = __builtin_fabsf(x);
    const bool inverted = a > 1.0f;
    if (inverted) a = 1.0f / a;  // 1/inf = 0
    const bool shifted = a > tan15;
    const float t = shifted ? (a - inv_sqrt3) / __builtin_fmaf(a, inv_sqrt3, 1.0f) : a;
    const float t2 = t * t;
    float p = -1.0f / 15.0f;
    p = __builtin_fmaf(p, t2, 1.0f / 13.0f);
    p = __builtin_fmaf(p, t2, -1.0f / 11.0f);
    p = __builtin_fmaf(p, t2, 1.0f / 9.0f);
    p = __builtin_fmaf(p, t2, -1.0f / 7.0f);
    p = __builtin_fmaf(p, t2, 1.0f / 5.0f);
    p = __builtin_fmaf(p, t2, -1.0f / 3.0f);
    float r = __builtin_fmaf(t * t2, p, t);  // t - t^3/3 + t^5/5 - ...
    if (shifted) r = (r + pi_6_lo) + pi_6_hi;
    if (inverted) r = (pi_2_lo - r) + pi_2_hi;
    return x < 0.0f ? -r : r;
}
