/*
 * mcx_math.h -- elementary float functions used by the transport kernel.
 *
 * Each function is a fixed sequence of correctly rounded binary32
 * operations (add, mul, fma, div, sqrt, rint, bit moves).  The kernel is
 * compiled with -ffp-contract=off and correctly rounded division/sqrt, so a
 * host build of the same sequence yields the same bits; the kernel's
 * photon trajectories are then reproducible on the CPU step for step.
 *
 * Methods (from the mathematics, no vendor library):
 *   exp:  Cody-Waite reduction by ln2 (split hi/lo), Taylor series to r^7,
 *         exponent assembly (two stages below 2^-125).
 *   log:  x = m 2^e with m in [1/sqrt2, sqrt2); log(m) = 2 atanh(s),
 *         s = (m-1)/(m+1), odd series to s^9.
 *   sincos(2 pi u): quadrant split of 4u, Taylor series to a^9 / a^10 on
 *         |a| <= pi/4.
 *   sincos(x): Cody-Waite reduction by pi/2 (three-part split), the same
 *         quadrant series.
 *
 * These stand in for CUDA's fast-math expf/logf/sincosf used by the
 * reference (src/mcx_core.cu:2140, :1900, :1932, :1958-1963, :1702).
 */
#ifndef MCX_AMD_MATH_H
#define MCX_AMD_MATH_H

#include <stdint.h>

#ifdef __HIPCC__
#define MCX_HD __host__ __device__ __forceinline__
#else
#define MCX_HD static inline
#include <math.h>
#include <string.h>
#endif

MCX_HD float mcx_bits2f(uint32_t u) { return __builtin_bit_cast(float, u); }
MCX_HD uint32_t mcx_f2bits(float f) { return __builtin_bit_cast(uint32_t, f); }

#define MCX_LN2_HI 0.693145751953125f
#define MCX_LN2_LO 1.428606765330187045e-06f

MCX_HD float mcx_expf(float x) {
    if (x < -104.0f) return 0.0f;
    if (x > 88.5f) return __builtin_inff();
    float kf = __builtin_rintf(x * 1.44269504088896341f);
    float r = __builtin_fmaf(-kf, MCX_LN2_HI, x);
    r = __builtin_fmaf(-kf, MCX_LN2_LO, r);
    float p = 1.0f / 5040.0f;
    p = __builtin_fmaf(p, r, 1.0f / 720.0f);
    p = __builtin_fmaf(p, r, 1.0f / 120.0f);
    p = __builtin_fmaf(p, r, 1.0f / 24.0f);
    p = __builtin_fmaf(p, r, 1.0f / 6.0f);
    p = __builtin_fmaf(p, r, 0.5f);
    p = __builtin_fmaf(p, r, 1.0f);
    p = __builtin_fmaf(p, r, 1.0f);
    int k = (int)kf;
    if (k < -125) {
        p = p * mcx_bits2f(0x1f800000u);
        k += 64;
    }
    return p * mcx_bits2f((uint32_t)(k + 127) << 23);
}

MCX_HD float mcx_logf(float x) {
    uint32_t ix = mcx_f2bits(x);
    int e = (int)(ix >> 23) - 127;
    float m = mcx_bits2f((ix & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e += 1;
    }
    float f = m - 1.0f;
    float s = f / (2.0f + f);
    float z = s * s;
    float p = 2.0f / 9.0f;
    p = __builtin_fmaf(p, z, 2.0f / 7.0f);
    p = __builtin_fmaf(p, z, 2.0f / 5.0f);
    p = __builtin_fmaf(p, z, 2.0f / 3.0f);
    float r = __builtin_fmaf(s * z, p, 2.0f * s);
    float fe = (float)e;
    return __builtin_fmaf(fe, MCX_LN2_HI, __builtin_fmaf(fe, MCX_LN2_LO, r));
}

/* sin and cos of a, |a| <= pi/4, rotated by the quadrant qi (0..3) */
MCX_HD void mcx_sincos_quadrant(float a, int qi, float* s, float* c) {
    float a2 = a * a;
    float ps = 1.0f / 362880.0f;
    ps = __builtin_fmaf(ps, a2, -1.0f / 5040.0f);
    ps = __builtin_fmaf(ps, a2, 1.0f / 120.0f);
    ps = __builtin_fmaf(ps, a2, -1.0f / 6.0f);
    float sn = __builtin_fmaf(a * a2, ps, a);
    float pc = -1.0f / 3628800.0f;
    pc = __builtin_fmaf(pc, a2, 1.0f / 40320.0f);
    pc = __builtin_fmaf(pc, a2, -1.0f / 720.0f);
    pc = __builtin_fmaf(pc, a2, 1.0f / 24.0f);
    pc = __builtin_fmaf(pc, a2, -0.5f);
    float cs = __builtin_fmaf(a2, pc, 1.0f);
    float ss = (qi & 1) ? cs : sn;
    float cc = (qi & 1) ? sn : cs;
    *s = (qi & 2) ? -ss : ss;
    *c = ((qi + 1) & 2) ? -cc : cc;
}

MCX_HD void mcx_sincos2pi(float u, float* s, float* c) {
    float x = u * 4.0f;
    float q = __builtin_rintf(x);
    float f = x - q;
    mcx_sincos_quadrant(f * 1.57079632679489662f, ((int)q) & 3, s, c);
}

/* sin and cos of any float x with |x| < 2^12 pi/2 (larger arguments keep
 * |s|, |c| <= 1 but lose accuracy): Cody-Waite reduction by pi/2 split in
 * three parts (8, 12 and 24 significant bits; q * hi and q * mid are exact
 * for |q| < 2^12), then the quadrant series. */
#define MCX_PIO2_HI 1.5703125f
#define MCX_PIO2_MID 4.837512969970703125e-04f
#define MCX_PIO2_LO 7.549790126404332e-08f

MCX_HD void mcx_sincosf(float x, float* s, float* c) {
    float q = __builtin_rintf(x * 0.636619772367581343f);
    float r = __builtin_fmaf(-q, MCX_PIO2_HI, x);
    r = __builtin_fmaf(-q, MCX_PIO2_MID, r);
    r = __builtin_fmaf(-q, MCX_PIO2_LO, r);
    mcx_sincos_quadrant(r, ((int)q) & 3, s, c);
}

/* sin(theta) from cos(theta) for theta in [0, pi] */
MCX_HD float mcx_sin_from_cos(float c) {
    return __builtin_sqrtf((1.0f - c) * (1.0f + c));
}

#endif
