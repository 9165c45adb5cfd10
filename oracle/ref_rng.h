// ORACLE (test infrastructure only): the random stream and the two
// transcendental functions the CPU re-statement uses.  Written separately
// from the product's device code; both must produce identical bits, which the
// tests check (Philox known-answer vectors, exp/log against FP64 libm, and
// GPU-vs-oracle sweeps).
//
// Philox4x32-10: Salmon et al., SC'11.  Round: (hi0, lo0) = M0*c0,
// (hi1, lo1) = M1*c2, c' = (hi1^c1^k0, lo1, hi0^c3^k1, lo0); key += Weyl.
// Counter layout: {id, step, (phase << 20) | draw, replica}, key = seed.
//
// Replaces include/tools/randomGenerator.h:31-94 (distribution shapes kept:
// randomUnit/randomReal uniform doubles, randomUnsigned uniform integers,
// geometric = floor(log(1-u)/log(1-p)) as in the reference's GPU formula).
#pragma once
#include <stdint.h>
#include <string.h>

namespace oracle {

inline void philox10(const uint32_t in[4], uint32_t key0, uint32_t key1, uint32_t out[4]) {
    uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ key0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ key1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
        key0 += 0x9E3779B9u;
        key1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// A draw sequence of one entity in one phase of one time step.
class Rng {
public:
    Rng(uint64_t seed, uint32_t id, uint32_t step, uint32_t phase, uint32_t replica = 0)
        : seed_(seed), id_(id), step_(step), phase_(phase), replica_(replica) {}
    void block(uint32_t out[4]) {
        uint32_t c[4] = {id_, step_, (phase_ << 20) | (n_ & 0xFFFFFu), replica_};
        philox10(c, (uint32_t)seed_, (uint32_t)(seed_ >> 32), out);
        ++n_;
    }
    double unit() {  // [0, 1), 53 bits
        uint32_t w[4];
        block(w);
        uint64_t bits = (((uint64_t)w[0] << 32) | w[1]) >> 11;
        return (double)bits / 9007199254740992.0;
    }
    double real(double max) { return max * unit(); }
    uint32_t below(uint32_t max) {  // [0, max); no draw when max == 0
        if (max == 0) return 0;
        uint32_t w[4];
        block(w);
        return (uint32_t)(((uint64_t)w[2] * max) >> 32);
    }
    uint32_t draws() const { return n_; }
    void seek(uint32_t n) { n_ = n; }  // continue the sequence at draw n

private:
    uint64_t seed_;
    uint32_t id_, step_, phase_, replica_;
    uint32_t n_ = 0;
};

enum : uint32_t {
    PHASE_INIT_DAYS = 1,
    PHASE_IMM_CATEGORY = 2,
    PHASE_MOVE = 3,
    PHASE_PLAN_SCHOOLWORK = 4,
    PHASE_TESTING = 5,
    PHASE_PROGRESSION = 6,
    PHASE_OTHER_DISEASE = 7,
    PHASE_INFECT = 8,
    PHASE_MUTATION = 16,
    PHASE_VACCINATION = 17,
    PHASE_TEST_FLAG = 18,
    PHASE_PLAN_UNDERAGE = 19,
    PHASE_DUMMY_MOVE = 20
};

inline float as_float(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}
inline uint64_t as_bits(double d) {
    uint64_t u;
    memcpy(&u, &d, 8);
    return u;
}
inline double as_double(uint64_t u) {
    double d;
    memcpy(&d, &u, 8);
    return d;
}

// exp for float: Cody-Waite reduction with ln2 = 0x1.62e4p-1 + 0x1.7f7d1cp-20,
// Taylor polynomial of exp(r) through r^8 written as 1 + (r + r^2 q(r)),
// scaling by 2^k through the exponent field.
inline float exp_f32(float x) {
    if (x != x) return x;
    if (x > 88.72283935546875f) return as_float(0x7f800000u);
    if (x < -103.972084045410156f) return 0.0f;
    float t = x * 0x1.715476p+0f + 0.5f;
    float kf = (float)(int)t;
    if (kf > t) kf = kf - 1.0f;  // floor
    int k = (int)kf;
    float r = (x - kf * 0x1.62e4p-1f) - kf * 0x1.7f7d1cp-20f;
    static const float c[7] = {2.48015873015873e-05f, 1.98412698412698e-04f, 1.38888888888889e-03f,
                               8.33333333333333e-03f, 4.16666666666667e-02f, 1.66666666666667e-01f, 0.5f};
    float q = c[0];
    for (int i = 1; i < 7; ++i) q = q * r + c[i];
    float p = 1.0f + (r + (r * r) * q);
    if (k > 127) return (p * as_float(254u << 23)) * 2.0f;
    if (k < -126) return (p * as_float((uint32_t)(k + 64 + 127) << 23)) * as_float((uint32_t)(127 - 64) << 23);
    return p * as_float((uint32_t)(k + 127) << 23);
}

// natural log for double: x = 2^e m, m in [sqrt(2)/2, sqrt(2)), f = m-1,
// s = f/(2+f); log m = f - f^2/2 + s (f^2/2 + sum_{k=1..12} 2 s^{2k}/(2k+1)).
inline double log_f64(double x) {
    if (x != x) return x;
    if (x < 0.0) return as_double(0x7ff8000000000000ull);
    if (x == 0.0) return as_double(0xfff0000000000000ull);
    uint64_t b = as_bits(x);
    if (b == 0x7ff0000000000000ull) return x;
    int e = (int)((b >> 52) & 0x7ff);
    if (e == 0) {
        x = x * 18014398509481984.0;
        b = as_bits(x);
        e = (int)((b >> 52) & 0x7ff) - 54;
    }
    e -= 1023;
    double m = as_double((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    if (m > 0x1.6a09e667f3bcdp+0) {
        m = m * 0.5;
        e += 1;
    }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double R = 2.0 / 25.0;
    for (int k = 11; k >= 1; --k) R = R * z + 2.0 / (double)(2 * k + 1);
    R = R * z;
    double hfsq = 0.5 * f * f;
    double de = (double)e;
    return de * 0x1.62e42fee00000p-1 + (f - (hfsq - (s * (hfsq + R) + de * 0x1.a39ef35793c76p-33)));
}

}  // namespace oracle
