// Project transcendental functions with bit-identical results on the host
// (x86-64 SSE2) and on gfx950.  The reference calls expf in the infection
// ratio (include/policies/infectionPolicies.h:545) and log in the geometric
// draw (include/tools/randomGenerator.h:211-216).  Vendor libm and the device
// math library do not round identically, so both sides use these functions,
// built only from IEEE +, -, *, / (no fused multiply-add: every translation
// unit is compiled with -ffp-contract=off) and exact bit manipulation.
//
// ps_expf: Cody & Waite (1980) additive range reduction x = k*ln2 + r with a
//   two-word ln2, |r| <= ln2/2, then the Taylor polynomial of exp(r) to r^8,
//   arranged as 1 + (r + r^2*q(r)) so that the final rounding dominates, and
//   an exact scaling by 2^k.  Error <= 1 ulp over the finite range.
// ps_log: reduction x = 2^e * m with m in [sqrt(2)/2, sqrt(2)),
//   f = m - 1, s = f/(2+f), log(m) = 2*atanh(s) = f - f^2/2 + s*(f^2/2 + R(s^2)),
//   R(z) = sum_{k>=1} 2 z^k/(2k+1) (Taylor coefficients, 12 terms), then
//   e*ln2 with a two-word ln2.  Error <= 1 ulp for normal and subnormal x > 0.
#pragma once
#include <stdint.h>

#include "ps_hd.h"

namespace ps {

PS_INLINE float bits_to_f32(uint32_t u) { return __builtin_bit_cast(float, u); }
PS_INLINE uint32_t f32_to_bits(float f) { return __builtin_bit_cast(uint32_t, f); }
PS_INLINE double bits_to_f64(uint64_t u) { return __builtin_bit_cast(double, u); }
PS_INLINE uint64_t f64_to_bits(double d) { return __builtin_bit_cast(uint64_t, d); }

// 2^e for e in [-126, 127]
PS_INLINE float pow2f(int e) { return bits_to_f32((uint32_t)(e + 127) << 23); }

PS_INLINE float ps_floorf(float x) {
    // floor for |x| < 2^23 via truncation toward zero and correction
    if (!(x > -8388608.0f && x < 8388608.0f)) return x;
    float t = (float)(int)x;
    return (t > x) ? t - 1.0f : t;
}

// exp(r) - 1 for |r| <= ln2/2: r + r^2*q(r) with
// q(r) = 1/2 + r/6 + r^2/24 + r^3/120 + r^4/720 + r^5/5040 + r^6/40320
PS_INLINE float expm1_poly(float r) {
    float q = 2.48015873015873e-05f;
    q = q * r + 1.98412698412698e-04f;
    q = q * r + 1.38888888888889e-03f;
    q = q * r + 8.33333333333333e-03f;
    q = q * r + 4.16666666666667e-02f;
    q = q * r + 1.66666666666667e-01f;
    q = q * r + 0.5f;
    return r + (r * r) * q;
}

PS_INLINE float ps_expf(float x) {
    if (x != x) return x;
    if (x > 88.72283935546875f) return bits_to_f32(0x7f800000u);
    if (x < -103.972084045410156f) return 0.0f;
    const float INV_LN2 = 0x1.715476p+0f;
    const float LN2_HI = 0x1.62e4p-1f;       // 16 significant bits: k*LN2_HI is exact
    const float LN2_LO = 0x1.7f7d1cp-20f;    // ln2 - LN2_HI rounded to float
    float kf = ps_floorf(x * INV_LN2 + 0.5f);
    int k = (int)kf;
    float r = (x - kf * LN2_HI) - kf * LN2_LO;
    float p = 1.0f + expm1_poly(r);
    if (k > 127) return (p * pow2f(127)) * 2.0f;
    if (k < -126) return (p * pow2f(k + 64)) * pow2f(-64);
    return p * pow2f(k);
}

PS_INLINE double ps_log(double x) {
    uint64_t b = f64_to_bits(x);
    if (x != x) return x;
    if (x < 0.0) return bits_to_f64(0x7ff8000000000000ull);
    if (x == 0.0) return bits_to_f64(0xfff0000000000000ull);
    if (b == 0x7ff0000000000000ull) return x;
    int e = (int)((b >> 52) & 0x7ff);
    if (e == 0) {  // subnormal: scale into the normal range
        x = x * 18014398509481984.0;  // 2^54
        b = f64_to_bits(x);
        e = (int)((b >> 52) & 0x7ff) - 54;
    }
    e -= 1023;
    double m = bits_to_f64((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    if (m > 0x1.6a09e667f3bcdp+0) {
        m = m * 0.5;
        e += 1;
    }
    const double LN2_HI = 0x1.62e42fee00000p-1;
    const double LN2_LO = 0x1.a39ef35793c76p-33;
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double R = 2.0 / 25.0;
    R = R * z + 2.0 / 23.0;
    R = R * z + 2.0 / 21.0;
    R = R * z + 2.0 / 19.0;
    R = R * z + 2.0 / 17.0;
    R = R * z + 2.0 / 15.0;
    R = R * z + 2.0 / 13.0;
    R = R * z + 2.0 / 11.0;
    R = R * z + 2.0 / 9.0;
    R = R * z + 2.0 / 7.0;
    R = R * z + 2.0 / 5.0;
    R = R * z + 2.0 / 3.0;
    R = R * z;
    double hfsq = 0.5 * f * f;
    double de = (double)e;
    return de * LN2_HI + (f - (hfsq - (s * (hfsq + R) + de * LN2_LO)));
}

// ps_expm1f: exp(x) - 1 without the cancellation of ps_expf(x) - 1 near 0.
// For |x| <= ln2/2 the reduction of ps_expf is the identity (k = 0, r = x)
// and the polynomial alone is exp(x) - 1 (Taylor to x^8: truncation below
// 2^-31 relative, then a few roundings); elsewhere |exp(x) - 1| >= 0.29 and
// ps_expf(x) - 1 loses at most a few ulp.
PS_INLINE float ps_expm1f(float x) {
    if (x >= -0.34657359f && x <= 0.34657359f) return expm1_poly(x);
    return ps_expf(x) - 1.0f;
}

}  // namespace ps
