/*
 * oracle_math.h -- TEST INFRASTRUCTURE ONLY (the oracle's float math).
 *
 * Elementary functions written as fixed sequences of IEEE-754 binary32
 * operations (add, multiply, fused multiply-add, exact division and square
 * root, rint, bit moves), so that any compiler that neither contracts nor
 * reassociates produces the same bits.  The methods come from the
 * mathematics, not from any vendor library:
 *
 *   exp(x):  x = k*ln2 + r (Cody-Waite split of ln2), |r| <= ln2/2,
 *            exp(r) by its Taylor series to r^7 (truncation < 6e-9 rel.),
 *            times 2^k by exponent assembly.
 *   log(x):  x = m * 2^e, m in [sqrt(1/2), sqrt(2)), f = m - 1,
 *            s = f/(2+f), log(1+f) = 2 atanh(s) = 2(s + s^3/3 + ... + s^9/9)
 *            (|s| <= 0.1716, truncation < 2e-9 rel.).
 *   sincos(2 pi u): u*4 = q + f, |f| <= 1/2, a = f*pi/2 in [-pi/4, pi/4],
 *            sin/cos by Taylor series to a^9 / a^10, quadrant q selects.
 *   sincos(x): q = rint(x * 2/pi), r = x - q*pi/2 with pi/2 split in three
 *            parts (Cody-Waite; 8 + 12 + 24 significant bits), the same
 *            series on r, quadrant q selects.
 *
 * The reference calls CUDA's fast-math expf/logf/sincosf/acosf
 * (src/mcx_core.cu:2140, :1900, :1932, :1958-1963, :1702, built with
 * -use_fast_math, src/Makefile:194); those are approximations whose bits
 * are unpinned.  The oracle only needs functions accurate to a few ulp:
 * tests/test_oracle_math.py checks these against libm.
 *
 * Where the reference computes sin(acos(c)) it is replaced by the identity
 * sqrt((1-c)(1+c)), and cos(acos(c)) by c itself.
 */
#ifndef MCX_ORACLE_MATH_H
#define MCX_ORACLE_MATH_H

#include <stdint.h>
#include <string.h>
#include <math.h>

static inline float om_bits2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline uint32_t om_f2bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

#define OM_LN2_HI 0.693145751953125f
#define OM_LN2_LO 1.428606765330187045e-06f

static inline float om_expf(float x) {
    if (x < -104.0f) return 0.0f;
    if (x > 88.5f) return INFINITY;
    float kf = rintf(x * 1.44269504088896341f);
    float r = fmaf(-kf, OM_LN2_HI, x);
    r = fmaf(-kf, OM_LN2_LO, r);
    float p = 1.0f / 5040.0f;
    p = fmaf(p, r, 1.0f / 720.0f);
    p = fmaf(p, r, 1.0f / 120.0f);
    p = fmaf(p, r, 1.0f / 24.0f);
    p = fmaf(p, r, 1.0f / 6.0f);
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    int k = (int)kf;
    if (k < -125) {
        p = p * om_bits2f(0x1f800000u); /* 2^-64 */
        k += 64;
    }
    return p * om_bits2f((uint32_t)(k + 127) << 23);
}

/* natural log for finite x > 0 in the normal range */
static inline float om_logf(float x) {
    uint32_t ix = om_f2bits(x);
    int e = (int)(ix >> 23) - 127;
    float m = om_bits2f((ix & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e += 1;
    }
    float f = m - 1.0f;
    float s = f / (2.0f + f);
    float z = s * s;
    float p = 2.0f / 9.0f;
    p = fmaf(p, z, 2.0f / 7.0f);
    p = fmaf(p, z, 2.0f / 5.0f);
    p = fmaf(p, z, 2.0f / 3.0f);
    float r = fmaf(s * z, p, 2.0f * s);
    float fe = (float)e;
    return fmaf(fe, OM_LN2_HI, fmaf(fe, OM_LN2_LO, r));
}

/* sin(a), cos(a) for |a| <= pi/4, then rotated by quadrant qi */
static inline void om_sincos_series(float a, int qi, float* s, float* c) {
    float a2 = a * a;
    float ps = 1.0f / 362880.0f;
    ps = fmaf(ps, a2, -1.0f / 5040.0f);
    ps = fmaf(ps, a2, 1.0f / 120.0f);
    ps = fmaf(ps, a2, -1.0f / 6.0f);
    float sn = fmaf(a * a2, ps, a);
    float pc = -1.0f / 3628800.0f;
    pc = fmaf(pc, a2, 1.0f / 40320.0f);
    pc = fmaf(pc, a2, -1.0f / 720.0f);
    pc = fmaf(pc, a2, 1.0f / 24.0f);
    pc = fmaf(pc, a2, -0.5f);
    float cs = fmaf(a2, pc, 1.0f);
    switch (qi) {
        case 0: *s = sn;  *c = cs;  break;
        case 1: *s = cs;  *c = -sn; break;
        case 2: *s = -sn; *c = -cs; break;
        default: *s = -cs; *c = sn; break;
    }
}

/* sin(2 pi u), cos(2 pi u) */
static inline void om_sincos2pi(float u, float* s, float* c) {
    float x = u * 4.0f;
    float q = rintf(x);
    float f = x - q;
    om_sincos_series(f * 1.57079632679489662f, ((int)q) & 3, s, c);
}

/* sin(x), cos(x): pi/2 = 1.5703125 + 4.837512969970703125e-4 + 7.549790126404332e-8 */
static inline void om_sincosf(float x, float* s, float* c) {
    const float q = rintf(x * 0.636619772367581343f);
    float r = fmaf(-q, 1.5703125f, x);
    r = fmaf(-q, 4.837512969970703125e-04f, r);
    r = fmaf(-q, 7.549790126404332e-08f, r);
    om_sincos_series(r, ((int)q) & 3, s, c);
}

/* sin of an angle whose cosine is c, c in [-1,1] */
static inline float om_sin_from_cos(float c) {
    return sqrtf((1.0f - c) * (1.0f + c));
}

#endif
