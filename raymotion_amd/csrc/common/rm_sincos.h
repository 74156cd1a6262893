#pragma once
// Deterministic single-precision sine/cosine for sampling angles.
//
// Host libm and the device math library round sinf/cosf differently in the last
// place, and a one-ulp change in a sampled direction can send a path through a
// different triangle. Both the HIP kernels and the host code therefore evaluate
// sin/cos with this one routine: Cody-Waite reduction by pi/2 (fdlibm's 33-bit
// split of pi/2) followed by the Taylor series of sin and cos on [-pi/4, pi/4]
// up to degree 17/18 (truncation error < 1e-19), all in IEEE double without
// fused operations, then rounded once to float. Every step is an exactly
// specified IEEE operation, so host and device results are bit-identical, and
// the double intermediate keeps the float result within one ulp of the
// correctly rounded value (checked against libm in the test suite).
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
    const double q = x * two_over_pi;
    double kf = q >= 0.0 ? double((long long)(q + 0.5)) : -double((long long)(0.5 - q));
    const long long k = (long long)kf;
    const double r = (x - kf * pio2_hi) - kf * pio2_lo;
    const double r2 = r * r;
    // sin(r) = r - r^3/3! + r^5/5! - ... ; cos(r) = 1 - r^2/2! + r^4/4! - ...
    double sp = 2.8114572543455206e-15;   // 1/17!
    sp = sp * r2 + -7.6471637318198164e-13;  // -1/15!
    sp = sp * r2 + 1.6059043836821613e-10;   // 1/13!
    sp = sp * r2 + -2.5052108385441720e-08;  // -1/11!
    sp = sp * r2 + 2.7557319223985893e-06;   // 1/9!
    sp = sp * r2 + -1.9841269841269841e-04;  // -1/7!
    sp = sp * r2 + 8.3333333333333333e-03;   // 1/5!
    sp = sp * r2 + -1.6666666666666667e-01;  // -1/3!
    const double sn = r + (r * r2) * sp;
    double cp = 1.5619206968586225e-16;   // 1/18!
    cp = cp * r2 + -4.7794773323873853e-14;  // -1/16!
    cp = cp * r2 + 1.1470745597729725e-11;   // 1/14!
    cp = cp * r2 + -2.0876756987868099e-09;  // -1/12!
    cp = cp * r2 + 2.7557319223985888e-07;   // 1/10!
    cp = cp * r2 + -2.4801587301587302e-05;  // -1/8!
    cp = cp * r2 + 1.3888888888888889e-03;   // 1/6!
    cp = cp * r2 + -4.1666666666666667e-02;  // -1/4!
    cp = cp * r2 + 0.5;                      // 1/2!
    const double cs = 1.0 - r2 * cp;
    double s, c;
    switch (k & 3) {
        case 0: s = sn; c = cs; break;
        case 1: s = cs; c = -sn; break;
        case 2: s = -sn; c = -cs; break;
        default: s = -cs; c = sn; break;
    }
    *s_out = float(s);
    *c_out = float(c);
}
