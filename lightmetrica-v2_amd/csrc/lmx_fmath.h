/*
 * lmx_fmath.h -- the FP32 sine, cosine, tangent, exponential, logarithm and
 * power of the path samplers, written once for the GPU kernels and for any
 * host code that must reproduce their arithmetic bit for bit.
 *
 * Every function is a fixed sequence of IEEE-754 single-precision
 * operations -- add, subtract, multiply, divide and fused multiply-add, each
 * correctly rounded, with no contraction, no reassociation and denormals
 * kept -- plus integer manipulation of the bit patterns, so any compiler and
 * processor that honours those rules (hipcc for gfx950 with
 * -ffp-contract=off and correctly rounded division, gcc or clang for x86-64
 * with -ffp-contract=off and no fast-math) computes the same bits.  The
 * vendor math libraries on either side are free to differ from each other by
 * an ulp, which is what this header exists to remove.
 *
 * Methods (textbook; no vendor library's constants or steps):
 *   sin, cos: Cody and Waite's additive range reduction (W. J. Cody and
 *     W. Waite, "Software Manual for the Elementary Functions", 1980):
 *     x = k pi/2 + r with pi/2 split into three floats and each step an FMA,
 *     |r| <= pi/4, then the Taylor series of sin r to r^9 and of cos r to
 *     r^10 (truncation below 0.05 ulp on |r| <= pi/4) and the quadrant k
 *     mod 4.  Valid for |x| <= 65536 (larger or non-finite x gives NaN).
 *   log: x = m 2^e with m in [sqrt(1/2), sqrt(2)), ln m = 2 atanh(s),
 *     s = (m - 1)/(m + 1), series to s^9 (truncation below 2e-9 relative),
 *     carried as a sum of two floats (the division's residual by FMA, e ln 2
 *     split so that its leading product is exact, a two-sum).
 *   exp: Cody-Waite reduction by ln 2 (two floats), Taylor series of e^r to
 *     r^7 on |r| <= ln2 / 2, scaling by 2^k in at most two exact steps.
 *   pow: x^y = exp(y ln x) for x >= 0 with the IEEE special cases below,
 *     y ln x kept to about 40 bits (the two-float log and the product's
 *     rounding error by FMA) so the error does not grow with |y ln x|.
 * Accuracy against double precision (tests/test_fmath.py): sin and cos
 * within 2 ulp on |x| <= 8, exp and log within 1 ulp, pow within 8 ulp
 * (99.99 % within 3) wherever the result is a normal float and |y| <= 1e4
 * (the samplers' ranges).
 */
#ifndef LMX_FMATH_H
#define LMX_FMATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#define LMX_FM static __host__ __device__ inline
#else
#define LMX_FM static inline
#endif

LMX_FM uint32_t lmx_f2u(float f)
{
    union { float f; uint32_t u; } c;
    c.f = f;
    return c.u;
}

LMX_FM float lmx_u2f(uint32_t u)
{
    union { float f; uint32_t u; } c;
    c.u = u;
    return c.f;
}

/* 2^n for n in [-126, 127], exactly */
LMX_FM float lmx_exp2i(int n) { return lmx_u2f((uint32_t)(n + 127) << 23); }

/* nearest integer (ties to even) of |x| < 2^22, as a float */
LMX_FM float lmx_rint_small(float x) { return (x + 0x1.8p23f) - 0x1.8p23f; }

/* sin x and cos x, |x| <= 65536 */
LMX_FM void lmx_sincosf(float x, float* s, float* c)
{
    const float ax = x < 0.0f ? -x : x;
    if (!(ax <= 65536.0f)) {               /* out of range, inf or NaN */
        *s = *c = lmx_u2f(0x7fc00000u);
        return;
    }
    const float k = lmx_rint_small(x * 0x1.45f306p-1f);            /* x 2/pi */
    float r = __builtin_fmaf(-k, 0x1.921fb6p+0f, x);                  /* pi/2, three parts */
    r = __builtin_fmaf(-k, -0x1.777a5cp-25f, r);
    r = __builtin_fmaf(-k, -0x1.ee59dap-50f, r);
    const float z = r * r;
    /* sin r = r + r^3 (-1/6 + z (1/120 + z (-1/5040 + z / 362880))) */
    const float ps = __builtin_fmaf(z, __builtin_fmaf(z, __builtin_fmaf(z, 2.75573192e-6f, -1.98412698e-4f),
                                                      8.33333333e-3f), -1.66666667e-1f);
    const float sr = __builtin_fmaf(r * z, ps, r);
    /* cos r = 1 - z/2 + z^2 (1/24 + z (-1/720 + z (1/40320 - z / 3628800))) */
    const float pc = __builtin_fmaf(z, __builtin_fmaf(z, __builtin_fmaf(z, -2.75573192e-7f, 2.48015873e-5f),
                                                      -1.38888889e-3f), 4.16666667e-2f);
    const float cr = __builtin_fmaf(z * z, pc, __builtin_fmaf(-0.5f, z, 1.0f));
    const int q = (int)k & 3;
    *s = q == 0 ? sr : (q == 1 ? cr : (q == 2 ? -sr : -cr));
    *c = q == 0 ? cr : (q == 1 ? -sr : (q == 2 ? -cr : sr));
}

LMX_FM float lmx_sinf(float x) { float s, c; lmx_sincosf(x, &s, &c); return s; }
LMX_FM float lmx_cosf(float x) { float s, c; lmx_sincosf(x, &s, &c); return c; }
LMX_FM float lmx_tanf(float x) { float s, c; lmx_sincosf(x, &s, &c); return s / c; }

/* natural logarithm of a positive finite x as an unevaluated sum hi + lo
 * (about 40 bits): s = f / (2 + f) keeps its division residual, the series
 * tail is added to that residual, and e ln 2 is split so that e times its
 * leading part is exact */
LMX_FM void lmx_log_ext(float x, float* hi, float* lo)
{
    int e = 0;
    if (x < 0x1p-126f) { x *= 0x1p23f; e = -23; }                    /* denormals */
    const uint32_t u = lmx_f2u(x);
    e += (int)((u >> 23) & 0xffu) - 127;
    float m = lmx_u2f((u & 0x007fffffu) | 0x3f800000u);               /* [1, 2) */
    if (m > 1.41421354f) { m *= 0.5f; e += 1; }                       /* [sqrt(1/2), sqrt(2)) */
    const float f = m - 1.0f;                                        /* exact */
    const float d = 2.0f + f;
    const float d_lo = f - (d - 2.0f);                               /* 2 + f = d + d_lo exactly */
    const float s = f / d;
    const float r = __builtin_fmaf(-s, d_lo, __builtin_fmaf(-s, d, f));
    const float s_lo = r / d;                                        /* f / (2 + f) = s + s_lo */
    const float z = s * s;
    /* ln m = 2 s (1 + z/3 + z^2/5 + z^3/7 + z^4/9) */
    const float p = __builtin_fmaf(z, __builtin_fmaf(z, __builtin_fmaf(z, 0.111111111f, 0.142857143f), 0.2f),
                                   0.333333333f);
    const float tail = __builtin_fmaf(2.0f * s * z, p, 2.0f * s_lo);
    const float fe = (float)e;
    const float a = fe * 0x1.62ep-1f;                                /* exact: 12 + 8 bits */
    const float b = 2.0f * s;
    const float sum = a + b;                                         /* two-sum of a and b */
    const float bv = sum - a;
    const float err = (a - (sum - bv)) + (b - bv);
    const float l = __builtin_fmaf(fe, 0x1.0bfbe8p-15f, err + tail);
    *hi = sum + l;
    *lo = l - (*hi - sum);
}

LMX_FM float lmx_logf_pos(float x)
{
    float hi, lo;
    lmx_log_ext(x, &hi, &lo);
    return hi + lo;
}

/* e^(t + tl), |tl| <= ulp(t) */
LMX_FM float lmx_exp_ext(float t, float tl)
{
    if (t != t) return t;
    if (t > 88.7228394f) return lmx_u2f(0x7f800000u);
    if (t < -103.972084f) return 0.0f;
    const float k = lmx_rint_small(t * 0x1.715476p+0f);              /* t / ln 2 */
    float r = __builtin_fmaf(-k, 0x1.62e430p-1f, t);                  /* ln 2, two parts; exact */
    r = __builtin_fmaf(-k, -0x1.05c610p-29f, r) + tl;
    /* e^r = 1 + r + r^2/2 + ... + r^7/5040 */
    float p = __builtin_fmaf(r, 1.98412698e-4f, 1.38888889e-3f);
    p = __builtin_fmaf(r, p, 8.33333333e-3f);
    p = __builtin_fmaf(r, p, 4.16666667e-2f);
    p = __builtin_fmaf(r, p, 1.66666667e-1f);
    p = __builtin_fmaf(r, p, 0.5f);
    p = __builtin_fmaf(r, p, 1.0f);
    p = __builtin_fmaf(r, p, 1.0f);
    const int n = (int)k;
    if (n > 127) return (p * lmx_exp2i(n - 64)) * 0x1p64f;
    if (n < -126) return (p * lmx_exp2i(n + 64)) * 0x1p-64f;         /* one rounding into the denormals */
    return p * lmx_exp2i(n);
}

LMX_FM float lmx_expf(float t) { return lmx_exp_ext(t, 0.0f); }

/* x^y for x >= 0 (x < 0 gives NaN): y = 0 or x = 1 -> 1; NaN -> NaN;
 * x = 0 -> 0 (y > 0) or +inf (y < 0); x = +inf -> +inf (y > 0) or 0.
 * y ln x is formed from the two-float logarithm with the product's own
 * rounding error (an FMA) kept, so the error of the result does not grow
 * with |y ln x| */
LMX_FM float lmx_powf(float x, float y)
{
    if (y == 0.0f || x == 1.0f) return 1.0f;
    if (x != x || y != y || x < 0.0f) return lmx_u2f(0x7fc00000u);
    if (x == 0.0f) return y > 0.0f ? 0.0f : lmx_u2f(0x7f800000u);
    if (x == lmx_u2f(0x7f800000u)) return y > 0.0f ? x : 0.0f;
    float hi, lo;
    lmx_log_ext(x, &hi, &lo);
    const float t = y * hi;
    const float tl = __builtin_fmaf(y, lo, __builtin_fmaf(y, hi, -t));
    return lmx_exp_ext(t, tl);
}

#endif
