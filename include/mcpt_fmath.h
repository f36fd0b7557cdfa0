/*
 * mcpt_fmath.h — portable single-precision elementary functions.
 *
 * Every function here is written with IEEE-754 binary32 +, -, *, /, sqrt,
 * floor and integer bit operations only, in a fixed evaluation order. Built
 * without floating-point contraction (-ffp-contract=off) and without
 * fast-math, any IEEE-conforming compiler produces identical bits for the
 * same input: g++ on x86-64 (SSE) and hipcc for gfx950 both do. This is what
 * lets the GPU kernels and the CPU oracle agree bit for bit on paths that
 * refract, scatter in a medium or look up the environment map.
 *
 * Methods (all published, coefficients fitted here):
 *   sin/cos   Cody & Waite (1980) argument reduction by pi/2 with a
 *             three-part constant (12 + 12 + 24 bits; exact products for
 *             |k| < 4096, i.e. |x| < 6400), then odd/even polynomials on
 *             [-pi/4, pi/4].
 *   atan      Cephes-style range reduction (tan(pi/8), tan(3pi/8)) and an
 *             odd polynomial on [0, tan(pi/8)]; atan2 by quadrant folding.
 *   asin/acos half-angle identity asin(x) = pi/2 - 2 asin(sqrt((1-x)/2))
 *             for |x| > 1/2, odd polynomial on [0, 1/2].
 *   log       x = m 2^e, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s),
 *             s = (m-1)/(m+1), ln 2 split 9 + 24 bits.
 *   exp       k = round(x log2 e), Cody-Waite r = x - k ln2, polynomial on
 *             [-ln2/2, ln2/2], scale by 2^k through the exponent field.
 *   pow       exp(y log x) for x > 0, without extra precision: the error
 *             grows with |y log x| (bound 2 + 2|y log x| ulp).
 * Polynomial coefficients are weighted least-squares fits at 4000 Chebyshev
 * nodes of the exact Taylor remainder, rounded to binary32.
 * Accuracy against libm (double-rounded reference) is checked by the test
 * suite: a few ulp on the documented domains.
 */
#ifndef MCPT_FMATH_H
#define MCPT_FMATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#define MCPT_HD __host__ __device__ inline
#else
#define MCPT_HD static inline
#endif

#define MCPT_PI_F 3.14159265358979323846f

MCPT_HD uint32_t mf_as_u32(float x) {
    uint32_t u;
    __builtin_memcpy(&u, &x, 4);
    return u;
}
MCPT_HD float mf_as_f32(uint32_t u) {
    float x;
    __builtin_memcpy(&x, &u, 4);
    return x;
}
MCPT_HD float mf_fabs(float x) { return mf_as_f32(mf_as_u32(x) & 0x7fffffffu); }
MCPT_HD float mf_copysign(float m, float s) {
    return mf_as_f32((mf_as_u32(m) & 0x7fffffffu) | (mf_as_u32(s) & 0x80000000u));
}
/* floor for |x| < 2^23 (larger values are already integers) */
MCPT_HD float mf_floor(float x) { return __builtin_floorf(x); }

/* ---------------------------------------------------------------- sin/cos */
MCPT_HD float mf_sin_poly(float r) {
    float z = r * r;
    return r + r * z * (-0.1666666567325592f + z * (0.008332819677889347f + z * -0.00019597770005930215f));
}
MCPT_HD float mf_cos_poly(float r) {
    float z = r * r;
    return 1.0f - 0.5f * z + z * z * (0.0416666641831398f + z * (-0.0013888375833630562f + z * 2.4558039513067342e-05f));
}
/* x = k pi/2 + r, |r| <= pi/4 (+rounding). Valid for |x| < 6400. */
MCPT_HD float mf_reduce_pio2(float x, int* quadrant) {
    float k = mf_floor(x * 0.6366197466850281f + 0.5f);
    float r = ((x - k * 1.57080078125f) - k * -4.453584551811218e-06f) - k * -8.705515752716053e-10f;
    *quadrant = (int)k;
    return r;
}
MCPT_HD float mf_sin(float x) {
    /* outside the reduction's exact range the result would be silently wrong:
       return NaN (quiet, propagates into the sample) instead */
    if (!(mf_fabs(x) < 6400.0f)) return __builtin_nanf("");
    int q;
    float r = mf_reduce_pio2(x, &q);
    switch (q & 3) {
        case 0: return mf_sin_poly(r);
        case 1: return mf_cos_poly(r);
        case 2: return -mf_sin_poly(r);
        default: return -mf_cos_poly(r);
    }
}
MCPT_HD float mf_cos(float x) {
    if (!(mf_fabs(x) < 6400.0f)) return __builtin_nanf("");
    int q;
    float r = mf_reduce_pio2(x, &q);
    switch (q & 3) {
        case 0: return mf_cos_poly(r);
        case 1: return -mf_sin_poly(r);
        case 2: return -mf_cos_poly(r);
        default: return mf_sin_poly(r);
    }
}

/* ------------------------------------------------------------- atan/atan2 */
MCPT_HD float mf_atan_pos(float x) { /* x >= 0 (or +inf) */
    float base = 0.0f;
    float t = x;
    if (x > 2.414213562373095f) {
        base = 1.5707963705062866f;
        t = -1.0f / x;
    } else if (x > 0.4142135623730950f) {
        base = 0.7853981852531433f;
        t = (x - 1.0f) / (x + 1.0f);
    }
    float z = t * t;
    float p = -0.3333331048488617f + z * (0.19992616772651672f + z * (-0.14041663706302643f + z * 0.08581631630659103f));
    return base + (t + t * z * p);
}
MCPT_HD float mf_atan(float x) { return mf_copysign(mf_atan_pos(mf_fabs(x)), x); }
MCPT_HD float mf_atan2(float y, float x) {
    float ax = mf_fabs(x), ay = mf_fabs(y);
    float a;
    if (ax == 0.0f && ay == 0.0f) a = 0.0f;
    else if (ay > ax) a = 1.5707963705062866f - mf_atan_pos(ax / ay);
    else a = mf_atan_pos(ay / ax);
    if (mf_as_u32(x) & 0x80000000u) a = 3.1415927410125732f - a;
    return (mf_as_u32(y) & 0x80000000u) ? -a : a;
}

/* ------------------------------------------------------------- asin/acos */
MCPT_HD float mf_asin_core(float s, float z) { /* s in [0, 1/2], z = s*s */
    float p = 0.1666666865348816f + z * (0.07499098777770996f + z * (0.04496048763394356f + z * (0.026792975142598152f + z * 0.03763545677065849f)));
    return s + s * z * p;
}
MCPT_HD float mf_asin(float x) {
    float a = mf_fabs(x);
    float r;
    if (a > 0.5f) {
        float z = (1.0f - a) * 0.5f;
        float s = __builtin_sqrtf(z);
        r = 1.5707963705062866f - 2.0f * mf_asin_core(s, z);
    } else {
        r = mf_asin_core(a, a * a);
    }
    return mf_copysign(r, x);
}
MCPT_HD float mf_acos(float x) {
    if (x < -0.5f) {
        float z = (1.0f + x) * 0.5f;
        float s = __builtin_sqrtf(z);
        return 3.1415927410125732f - 2.0f * mf_asin_core(s, z);
    }
    if (x > 0.5f) {
        float z = (1.0f - x) * 0.5f;
        float s = __builtin_sqrtf(z);
        return 2.0f * mf_asin_core(s, z);
    }
    float a = mf_fabs(x);
    float r = mf_asin_core(a, a * a);
    return 1.5707963705062866f - mf_copysign(r, x);
}

/* -------------------------------------------------------------------- log */
MCPT_HD float mf_log(float x) {
    if (!(x > 0.0f)) return (x == 0.0f) ? -__builtin_inff() : (x - x) / (x - x);
    if (x == __builtin_inff()) return x;
    int e = 0;
    if (x < 1.17549435e-38f) { /* subnormal: scale into the normal range */
        x = x * 8388608.0f;
        e = -23;
    }
    uint32_t bits = mf_as_u32(x);
    e += (int)(bits >> 23) - 127;
    float m = mf_as_f32((bits & 0x007fffffu) | 0x3f800000u);
    if (m > 1.4142135623730951f) {
        m = m * 0.5f;
        e += 1;
    }
    float f = m - 1.0f;
    float s = f / (2.0f + f);
    float z = s * s;
    float p = 0.6666666865348816f + z * (0.40000101923942566f + z * (0.28552407026290894f + z * 0.2329765260219574f));
    float lm = 2.0f * s + s * z * p;
    float ef = (float)e;
    return ef * 0.693359375f + (lm + ef * -0.00021219444170128554f);
}

/* -------------------------------------------------------------------- exp */
MCPT_HD float mf_exp(float x) {
    if (x != x) return x;
    if (x > 88.72283935546875f) return __builtin_inff();
    if (x < -103.97207641601562f) return 0.0f;
    float k = mf_floor(x * 1.4426950216293335f + 0.5f);
    float r = (x - k * 0.693359375f) - k * -0.00021219444170128554f;
    float p = 0.5f + r * (0.16666577756404877f + r * (0.041666943579912186f + r * (0.00836312584578991f + r * 0.0013893831055611372f)));
    float y = 1.0f + (r + r * r * p);
    int ki = (int)k;
    if (ki > 127) {
        y = y * 2.0f;
        ki -= 1;
    }
    if (ki < -126) {
        y = y * 5.421010862427522e-20f; /* 2^-64 */
        ki += 64;
    }
    return y * mf_as_f32((uint32_t)(ki + 127) << 23);
}

/* -------------------------------------------------------------------- pow */
/* x >= 0 only (every caller raises a non-negative base). */
MCPT_HD float mf_pow(float x, float y) {
    if (y == 0.0f || x == 1.0f) return 1.0f;
    if (x == 0.0f) return (y > 0.0f) ? 0.0f : __builtin_inff();
    if (!(x > 0.0f)) return (x - x) / (x - x);
    return mf_exp(y * mf_log(x));
}

#endif /* MCPT_FMATH_H */
