#pragma once
// Deterministic single-precision sine/cosine for sampling angles.
//
// Host libm and the device math library round sinf/cosf differently in the last
// place, and a one-ulp change in a sampled direction can send a path through a
// different triangle. Both the HIP kernels and the host code therefore evaluate
// sin/cos with this one routine: Cody-Waite reduction by pi/2 (fdlibm's 33-bit
// split of pi/2; the quotient rounded to nearest) followed by the Taylor series
// of sin and cos on [-pi/4, pi/4] up to degree 13/14 (truncation error below
// 3e-14 relative), in IEEE double with explicit fused multiply-adds (Horner),
// then rounded once to float. Every step is an exactly specified IEEE
// operation (fma is correctly rounded on host and device alike), so host and
// device results are bit-identical, and the double intermediate keeps the
// float result within one ulp of the correctly rounded value (checked against
// libm in the test suite).
//
// Valid for |x| < 2^20; sampling uses x in [0, 2*pi).
#if defined(__HIPCC__)
#define RM_SINCOS_HD __host__ __device__
#else
#define RM_SINCOS_HD
#endif

RM_SINCOS_HD inline void rm_sincosf(float xf, float* s_out, float* c_out) {
    const double x = double(xf);
    const double two_over_pi = 0.63661977236758134308;
    const double pio2_hi = 1.57079632673412561417e+00;  // first 33 bits of pi/2
    const double pio2_lo = 6.07710050650619224932e-11;  // pi/2 - pio2_hi
    const double kf = __builtin_rint(x * two_over_pi);
    const int k = int(kf);
    const double r = (x - kf * pio2_hi) - kf * pio2_lo;
    const double r2 = r * r;
    // sin(r) = r - r^3/3! + r^5/5! - ... ; cos(r) = 1 - r^2/2! + r^4/4! - ...
    double sp = 1.6059043836821613e-10;                 // 1/13!
    sp = __builtin_fma(sp, r2, -2.5052108385441720e-08);  // -1/11!
    sp = __builtin_fma(sp, r2, 2.7557319223985893e-06);   // 1/9!
    sp = __builtin_fma(sp, r2, -1.9841269841269841e-04);  // -1/7!
    sp = __builtin_fma(sp, r2, 8.3333333333333333e-03);   // 1/5!
    sp = __builtin_fma(sp, r2, -1.6666666666666667e-01);  // -1/3!
    const double sn = __builtin_fma(r * r2, sp, r);
    double cp = -4.7794773323873853e-14;                // -1/16!
    cp = __builtin_fma(cp, r2, 1.1470745597729725e-11);   // 1/14!
    cp = __builtin_fma(cp, r2, -2.0876756987868099e-09);  // -1/12!
    cp = __builtin_fma(cp, r2, 2.7557319223985888e-07);   // 1/10!
    cp = __builtin_fma(cp, r2, -2.4801587301587302e-05);  // -1/8!
    cp = __builtin_fma(cp, r2, 1.3888888888888889e-03);   // 1/6!
    cp = __builtin_fma(cp, r2, -4.1666666666666667e-02);  // -1/4!
    cp = __builtin_fma(cp, r2, 0.5);                      // 1/2!
    const double cs = __builtin_fma(-r2, cp, 1.0);
    const bool odd = (k & 1) != 0, neg_s = (k & 2) != 0, neg_c = ((k + 1) & 2) != 0;
    const double s = odd ? cs : sn, c = odd ? sn : cs;
    *s_out = float(neg_s ? -s : s);
    *c_out = float(neg_c ? -c : c);
}
