/*
 * esth_math.h - deterministic single-precision transcendentals.
 *
 * Every function here is a fixed sequence of IEEE-754 binary32 operations
 * (+, -, *, fused multiply-add, correctly rounded division and square root,
 * exact bit manipulation).  Compiled without floating-point contraction, the
 * same source therefore yields bit-identical results on gfx950 and on an x86
 * host.  That is what lets the GPU step be compared bit for bit with a CPU
 * restatement of the filter, including the model likelihoods.
 *
 * The reference evaluates these functions with OpenCL built-ins under
 * -cl-fast-relaxed-math (src/partfilt/opencl/pf_generic.c:1166); their exact
 * results belong to whatever OpenCL runtime ran it.  The engine pins its own
 * functions instead, and bounds their error against double-precision libm.
 *
 * Methods (all textbook):
 *   exp   : Cody-Waite reduction x = n ln2 + r, |r| <= ln2/2, Taylor degree 7.
 *   log   : x = m 2^e, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh((m-1)/(m+1))
 *           by its odd series to f^11.
 *   sincos: Cody-Waite reduction by pi/2 with a three-part constant, Taylor
 *           polynomials on [-pi/4, pi/4], quadrant selection.
 *   atan2 : octant reduction t = min/max in [0,1], Chebyshev-interpolated
 *           polynomial of atan(t)/t in t^2 (coefficients from gen_atan_poly.py).
 *
 * Domain notes: sincos is accurate for |x| < 1e5 (Cody-Waite range); exp
 * returns exactly 0 below -87.33 (no subnormal results) and +inf above 88.72.
 */
#ifndef ESTH_MATH_H
#define ESTH_MATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define ESTH_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#include <string.h>
#define ESTH_HD static inline
#endif

ESTH_HD uint32_t esth_f2u(float f) {
#if defined(__HIPCC__)
    return __builtin_bit_cast(uint32_t, f);
#else
    uint32_t u; memcpy(&u, &f, 4); return u;
#endif
}

ESTH_HD float esth_u2f(uint32_t u) {
#if defined(__HIPCC__)
    return __builtin_bit_cast(float, u);
#else
    float f; memcpy(&f, &u, 4); return f;
#endif
}

ESTH_HD float esth_fma(float a, float b, float c) { return fmaf(a, b, c); }

/* round to nearest integer, ties away from zero is not needed: rintf is ties-to-even on both sides */
ESTH_HD float esth_rint(float x) { return rintf(x); }

/* 2^n for integer n in [-126, 127] built from the exponent field (exact). */
ESTH_HD float esth_pow2i(int n) { return esth_u2f((uint32_t)(n + 127) << 23); }

#define ESTH_LN2_HI   0.693145751953125f          /* 0x3f317200: ln2 to 16 bits */
#define ESTH_LN2_LO   1.428606765330187e-06f      /* ln2 - LN2_HI, rounded to binary32 */
#define ESTH_INV_LN2  1.44269504088896340736f
#define ESTH_PI       3.14159265358979323846f
#define ESTH_PI_2     1.57079632679489661923f
#define ESTH_2PI      6.28318530717958647692f
/* pi/2 split into three binary32 parts: PIO2_1 + PIO2_2 + PIO2_3 ~ pi/2 */
#define ESTH_PIO2_1   1.5707963705062866211f       /* 0x3fc90fdb */
#define ESTH_PIO2_2  -4.3711388286737928865e-08f   /* 0xb33bbd2e */
#define ESTH_PIO2_3  -1.7151245100058819e-15f      /* 0xa6f72ced */
#define ESTH_2_OVER_PI 0.63661977236758134308f

/* e^x.  exp(x) = 2^n * e^r with n = rint(x/ln2), r = x - n ln2 (two-part ln2).
 * Branch-free: the polynomial runs on x clamped to [-87.33, 88.72] and the
 * out-of-range and NaN results are selected at the end; 2^n is applied as
 * 2^(n>>1) * 2^(n - (n>>1)), both factors normal, the first product exact, so
 * the one rounding is that of p * 2^n. */
ESTH_HD float esth_expf(float x) {
    const float xc = fminf(fmaxf(x, -87.33f), 88.72f);
    float n = esth_rint(xc * ESTH_INV_LN2);
    float r = esth_fma(-n, ESTH_LN2_HI, xc);
    r = esth_fma(-n, ESTH_LN2_LO, r);
    /* Taylor degree 7: 1 + r + r^2/2! + ... + r^7/7! */
    float p = 1.98412698412698412698e-04f;            /* 1/5040 */
    p = esth_fma(p, r, 1.38888888888888888889e-03f);  /* 1/720 */
    p = esth_fma(p, r, 8.33333333333333333333e-03f);  /* 1/120 */
    p = esth_fma(p, r, 4.16666666666666666667e-02f);  /* 1/24 */
    p = esth_fma(p, r, 1.66666666666666666667e-01f);  /* 1/6 */
    p = esth_fma(p, r, 0.5f);
    p = esth_fma(p, r, 1.0f);
    p = esth_fma(p, r, 1.0f);
    const int ni = (int)n;
    const int n1 = ni >> 1;
    float y = (p * esth_pow2i(n1)) * esth_pow2i(ni - n1);
    y = (x < -87.33f) ? 0.0f : y;
    y = (x > 88.72f) ? esth_u2f(0x7f800000u) : y;
    return (x != x) ? x : y;
}

/* the shift of a log-domain normalisation exp(x - shift) by the maximum m:
 * m itself, or 0 when every x is -inf (weights 0 rather than NaN) */
ESTH_HD float esth_max_shift(float m) { return m == -INFINITY ? 0.0f : m; }

/* natural log for x > 0 (x <= 0 returns -inf / NaN). */
ESTH_HD float esth_logf(float x) {
    if (!(x > 0.0f)) return (x == 0.0f) ? esth_u2f(0xff800000u) : esth_u2f(0x7fc00000u);
    if (x == esth_u2f(0x7f800000u)) return x;
    int e = 0;
    if (x < 1.17549435e-38f) { x *= 8388608.0f; e = -23; } /* subnormal input: scale by 2^23 */
    uint32_t b = esth_f2u(x);
    e += (int)((b >> 23) & 0xffu) - 127;
    float m = esth_u2f((b & 0x007fffffu) | 0x3f800000u); /* m in [1,2) */
    if (m > 1.41421356237309504880f) { m *= 0.5f; e += 1; }
    float f = (m - 1.0f) / (m + 1.0f);
    float s = f * f;
    float p = 9.09090909090909090909e-02f;            /* 1/11 */
    p = esth_fma(p, s, 1.11111111111111111111e-01f);  /* 1/9 */
    p = esth_fma(p, s, 1.42857142857142857143e-01f);  /* 1/7 */
    p = esth_fma(p, s, 2.0e-01f);                     /* 1/5 */
    p = esth_fma(p, s, 3.33333333333333333333e-01f);  /* 1/3 */
    float f2 = f + f;
    float lm = esth_fma(f2 * s, p, f2);               /* 2f + 2f s P(s) */
    float ef = (float)e;
    return esth_fma(ef, ESTH_LN2_HI, esth_fma(ef, ESTH_LN2_LO, lm));
}

/* sin and cos of x. */
ESTH_HD void esth_sincosf(float x, float* sp, float* cp) {
    float k = esth_rint(x * ESTH_2_OVER_PI);
    float r = esth_fma(-k, ESTH_PIO2_1, x);
    r = esth_fma(-k, ESTH_PIO2_2, r);
    r = esth_fma(-k, ESTH_PIO2_3, r);
    float r2 = r * r;
    /* sin r = r + r^3 (-1/6 + r^2 (1/120 + r^2 (-1/5040 + r^2/362880))) */
    float ps = 2.75573192239858906526e-06f;
    ps = esth_fma(ps, r2, -1.98412698412698412698e-04f);
    ps = esth_fma(ps, r2, 8.33333333333333333333e-03f);
    ps = esth_fma(ps, r2, -1.66666666666666666667e-01f);
    float s = esth_fma(r * r2, ps, r);
    /* cos r = 1 - r^2/2 + r^4 (1/24 + r^2 (-1/720 + r^2 (1/40320 - r^2/3628800))) */
    float pc = -2.75573192239858906526e-07f;
    pc = esth_fma(pc, r2, 2.48015873015873015873e-05f);
    pc = esth_fma(pc, r2, -1.38888888888888888889e-03f);
    pc = esth_fma(pc, r2, 4.16666666666666666667e-02f);
    float hr = 0.5f * r2;
    float w = 1.0f - hr;
    float c = w + ((((1.0f - w) - hr)) + (r2 * r2) * pc);
    int q = ((int)k) & 3;
    float so, co;
    if (q == 0)      { so = s;  co = c;  }
    else if (q == 1) { so = c;  co = -s; }
    else if (q == 2) { so = -s; co = -c; }
    else             { so = -c; co = s;  }
    *sp = so;
    *cp = co;
}

#define ESTH_ATAN_C0 1.0f
#define ESTH_ATAN_C1 -0.3333328366279602f
#define ESTH_ATAN_C2 0.19998230040073395f
#define ESTH_ATAN_C3 -0.14261573553085327f
#define ESTH_ATAN_C4 0.10940198600292206f
#define ESTH_ATAN_C5 -0.08372063934803009f
#define ESTH_ATAN_C6 0.05746355652809143f
#define ESTH_ATAN_C7 -0.030717508867383003f
#define ESTH_ATAN_C8 0.010680719278752804f
#define ESTH_ATAN_C9 -0.0017437011701986194f

/* atan(t) for t in [0, 1]. */
ESTH_HD float esth_atanf_unit(float t) {
    float s = t * t;
    float p = ESTH_ATAN_C9;
    p = esth_fma(p, s, ESTH_ATAN_C8);
    p = esth_fma(p, s, ESTH_ATAN_C7);
    p = esth_fma(p, s, ESTH_ATAN_C6);
    p = esth_fma(p, s, ESTH_ATAN_C5);
    p = esth_fma(p, s, ESTH_ATAN_C4);
    p = esth_fma(p, s, ESTH_ATAN_C3);
    p = esth_fma(p, s, ESTH_ATAN_C2);
    p = esth_fma(p, s, ESTH_ATAN_C1);
    p = esth_fma(p, s, ESTH_ATAN_C0);
    return t * p;
}

/* atan2(y, x) in [-pi, pi]; atan2(0, 0) = 0. */
ESTH_HD float esth_atan2f(float y, float x) {
    float ax = fabsf(x), ay = fabsf(y);
    float mx = ax > ay ? ax : ay;
    float mn = ax > ay ? ay : ax;
    float t = mn / mx;
    t = (mx == 0.0f) ? 0.0f : t;   /* atan2(0, 0) = 0 without a branch */
    float a = esth_atanf_unit(t);
    if (ay > ax) a = ESTH_PI_2 - a;
    if (x < 0.0f) a = ESTH_PI - a;
    return (y < 0.0f) ? -a : a;
}

/* sqrt(1 + s) on s in [0, 1]: degree-7 fit (gen_atan_poly.py), max relative
 * error 1.2e-7 */
#define ESTH_HYP_C0 1.0f                     /* 0x3f800000 */
#define ESTH_HYP_C1 0.4999963343143463f      /* 0x3effff85 */
#define ESTH_HYP_C2 -0.12491840124130249f    /* 0xbdffd538 */
#define ESTH_HYP_C3 0.061791423708200455f    /* 0x3d7d1901 */
#define ESTH_HYP_C4 -0.03592449799180031f    /* 0xbd132591 */
#define ESTH_HYP_C5 0.019269991666078568f    /* 0x3c9ddc1a */
#define ESTH_HYP_C6 -0.007379245478659868f   /* 0xbbf1cd99 */
#define ESTH_HYP_C7 0.001377946580760181f    /* 0x3ab49c37 */

/* 1/x for normal x > 0 without a divide: an integer first guess (relative
 * error below 0.051 over every binade) and three Newton steps y += y (1 - x y)
 * in fused multiply-adds, each squaring the error; within 1 ulp. */
ESTH_HD float esth_rcpf(float x) {
    float y = esth_u2f(0x7ef311c3u - esth_f2u(x));
    float e = esth_fma(-x, y, 1.0f);
    y = esth_fma(y, e, y);
    e = esth_fma(-x, y, 1.0f);
    y = esth_fma(y, e, y);
    e = esth_fma(-x, y, 1.0f);
    return esth_fma(y, e, y);
}

/* Length r and |angle| of the vector (u, v): r = max * sqrt(1 + t^2) and
 * |atan2(v, u)| in [0, pi] from one ratio t = min/max in [0, 1] (one
 * reciprocal, one square shared by both polynomials).  r is within about
 * 3 ulp of hypot(u, v), the angle within 1.2e-7 rad.  Only multiplies, fused
 * multiply-adds, min/max, compares and an integer subtract: no divide and no
 * square root, whose hardware sequences are the costly part of a likelihood. */
ESTH_HD void esth_polar_abs(float u, float v, float* r, float* ang) {
    const float au = fabsf(u), av = fabsf(v);
    const float mx = fmaxf(au, av), mn = fminf(au, av);
    const float t = mn * esth_rcpf(fmaxf(mx, 1e-30f));
    const float s = t * t;
    float a = esth_atanf_unit(t);
    a = (av > au) ? ESTH_PI_2 - a : a;
    *ang = (u < 0.0f) ? ESTH_PI - a : a;
    float q = ESTH_HYP_C7;
    q = esth_fma(q, s, ESTH_HYP_C6);
    q = esth_fma(q, s, ESTH_HYP_C5);
    q = esth_fma(q, s, ESTH_HYP_C4);
    q = esth_fma(q, s, ESTH_HYP_C3);
    q = esth_fma(q, s, ESTH_HYP_C2);
    q = esth_fma(q, s, ESTH_HYP_C1);
    q = esth_fma(q, s, ESTH_HYP_C0);
    *r = mx * q;
}

/* wrap an angle to (-pi, pi] using a 2-part 2*pi */
ESTH_HD float esth_wrap_pi(float a) {
    float k = esth_rint(a * (1.0f / ESTH_2PI));
    float r = esth_fma(-k, 6.28318548202514648438f, a);   /* 0x40c90fdb */
    r = esth_fma(-k, -1.7484555314695172e-07f, r);       /* 2pi - 0x40c90fdb (0xb43bbd2e) */
    if (r <= -ESTH_PI) r += ESTH_2PI;
    if (r > ESTH_PI) r -= ESTH_2PI;
    return r;
}

#endif /* ESTH_MATH_H */
