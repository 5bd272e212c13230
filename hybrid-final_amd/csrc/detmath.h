/*
 * detmath.h - deterministic elementary functions for the PIC-MCC kernels.
 *
 * The collision physics needs log10, exp, sin, cos, tan, atan and pow.  The
 * platform libraries (glibc on the host, the ROCm device library on gfx950)
 * agree only to about one ulp, which is enough to flip a rare
 * "R <= sum sigma" decision and break bit-exact event counts between the GPU
 * path and its CPU check.  Every function here is built only from IEEE-754
 * double +, -, *, / and exact bit manipulation, evaluated in a fixed order,
 * so any conforming compiler produces the same bits on any target as long as
 * floating-point contraction is disabled (-ffp-contract=off).
 *
 * Methods (all classical, written from the mathematics):
 *   exp  : Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, Taylor series
 *          of exp(r) to degree 14, scaling by 2^k through the exponent bits.
 *   log  : x = m 2^e with m in [sqrt(1/2), sqrt(2)), s = (m-1)/(m+1),
 *          log m = 2 atanh(s) by its odd power series to s^23.
 *   sin/cos : reduction by pi/2 with a three-part Cody-Waite constant,
 *          Taylor series on |r| <= pi/4, quadrant selection.
 *   atan : atan(x) = pi/2 - atan(1/x) for |x| > 1, then two half-angle
 *          steps atan(x) = 2 atan(x / (1 + sqrt(1 + x^2))) to |x| <= 0.2,
 *          then the odd power series to x^29.
 * Accuracy is better than 2e-15 relative over the ranges the kernels use.
 */
#ifndef PIC_DETMATH_H
#define PIC_DETMATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#define DM_FN static inline __host__ __device__
#else
#define DM_FN static inline
#endif

DM_FN uint64_t dm_bits(double x) {
    union { double d; uint64_t u; } v;
    v.d = x;
    return v.u;
}
DM_FN double dm_from_bits(uint64_t u) {
    union { double d; uint64_t u; } v;
    v.u = u;
    return v.d;
}

/* 2^k for integer k in the normal range, exact. */
DM_FN double dm_pow2i(int k) {
    if (k < -1022) k = -1022;
    if (k > 1023) k = 1023;
    return dm_from_bits((uint64_t)(k + 1023) << 52);
}

/* round-half-away-from-zero to an integer-valued double (exact for |x|<2^51) */
DM_FN double dm_rint(double x) {
    double t = x < 0.0 ? -x : x;
    double r = (double)(int64_t)(t + 0.5);
    return x < 0.0 ? -r : r;
}

#define DM_LN2_HI 6.93147180369123816490e-01  /* high 32 bits of ln 2 */
#define DM_LN2_LO 1.90821492927058770002e-10  /* ln 2 - DM_LN2_HI */
#define DM_INV_LN2 1.44269504088896338700e+00
#define DM_INV_LN10 4.34294481903251816668e-01
#define DM_PI 3.14159265358979311600e+00
#define DM_PIO2 1.57079632679489655800e+00

DM_FN double dm_exp(double x) {
    if (x > 709.0) x = 709.0;
    if (x < -745.0) return 0.0;
    double k = dm_rint(x * DM_INV_LN2);
    double r = x - k * DM_LN2_HI;
    r = r - k * DM_LN2_LO;
    /* Taylor series of exp(r), Horner form, degree 14 */
    double p = 1.0 / 87178291200.0;            /* 1/14! */
    p = p * r + 1.0 / 6227020800.0;            /* 1/13! */
    p = p * r + 1.0 / 479001600.0;
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    int ki = (int)k;
    if (ki < -1022) {               /* gradual underflow: scale in two steps */
        return p * dm_pow2i(ki + 600) * dm_pow2i(-600);
    }
    return p * dm_pow2i(ki);
}

/* natural log for x > 0 (returns -745 for x <= 0 so callers never see NaN) */
DM_FN double dm_log(double x) {
    if (!(x > 0.0)) return -745.0;
    uint64_t u = dm_bits(x);
    int e = (int)((u >> 52) & 0x7ff);
    if (e == 0) {                   /* subnormal: renormalise */
        x = x * 4503599627370496.0; /* 2^52 */
        u = dm_bits(x);
        e = (int)((u >> 52) & 0x7ff) - 52;
    }
    e -= 1023;
    double m = dm_from_bits((u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    if (m > 1.41421356237309514547) { m = m * 0.5; e += 1; }
    double s = (m - 1.0) / (m + 1.0);
    double s2 = s * s;
    double p = 1.0 / 23.0;
    p = p * s2 + 1.0 / 21.0;
    p = p * s2 + 1.0 / 19.0;
    p = p * s2 + 1.0 / 17.0;
    p = p * s2 + 1.0 / 15.0;
    p = p * s2 + 1.0 / 13.0;
    p = p * s2 + 1.0 / 11.0;
    p = p * s2 + 1.0 / 9.0;
    p = p * s2 + 1.0 / 7.0;
    p = p * s2 + 1.0 / 5.0;
    p = p * s2 + 1.0 / 3.0;
    p = p * s2 + 1.0;
    double lm = 2.0 * s * p;
    double de = (double)e;
    return de * DM_LN2_HI + (de * DM_LN2_LO + lm);
}

DM_FN double dm_log10(double x) { return dm_log(x) * DM_INV_LN10; }

/* a^b for a > 0 */
DM_FN double dm_pow(double a, double b) { return dm_exp(b * dm_log(a)); }

/* sin(r), cos(r) Taylor series for |r| <= pi/4 */
DM_FN double dm_sin_kernel(double r) {
    double r2 = r * r;
    double p = -1.0 / 1307674368000.0;          /* -1/15! */
    p = p * r2 + 1.0 / 6227020800.0;            /*  1/13! */
    p = p * r2 - 1.0 / 39916800.0;              /* -1/11! */
    p = p * r2 + 1.0 / 362880.0;                /*  1/9!  */
    p = p * r2 - 1.0 / 5040.0;
    p = p * r2 + 1.0 / 120.0;
    p = p * r2 - 1.0 / 6.0;
    p = p * r2;
    return r + r * p;
}
DM_FN double dm_cos_kernel(double r) {
    double r2 = r * r;
    double p = 1.0 / 20922789888000.0;          /*  1/16! */
    p = p * r2 - 1.0 / 87178291200.0;           /* -1/14! */
    p = p * r2 + 1.0 / 479001600.0;             /*  1/12! */
    p = p * r2 - 1.0 / 3628800.0;               /* -1/10! */
    p = p * r2 + 1.0 / 40320.0;
    p = p * r2 - 1.0 / 720.0;
    p = p * r2 + 1.0 / 24.0;
    p = p * r2 - 0.5;
    p = p * r2;
    return 1.0 + p;
}

/* pi/2 split in three parts (33 + 33 + 53 bits) for Cody-Waite reduction */
#define DM_PIO2_1 1.57079632673412561417e+00
#define DM_PIO2_2 6.07710050630396597660e-11
#define DM_PIO2_3 2.02226624879595063154e-21

DM_FN double dm_reduce_pio2(double x, int *quadrant) {
    double k = dm_rint(x * 6.36619772367581382433e-01); /* 2/pi */
    double r = x - k * DM_PIO2_1;
    r = r - k * DM_PIO2_2;
    r = r - k * DM_PIO2_3;
    int64_t ki = (int64_t)k;
    *quadrant = (int)(ki & 3);
    return r;
}

DM_FN double dm_sin(double x) {
    int q;
    double r = dm_reduce_pio2(x, &q);
    switch (q) {
    case 0: return dm_sin_kernel(r);
    case 1: return dm_cos_kernel(r);
    case 2: return -dm_sin_kernel(r);
    default: return -dm_cos_kernel(r);
    }
}
DM_FN double dm_cos(double x) {
    int q;
    double r = dm_reduce_pio2(x, &q);
    switch (q) {
    case 0: return dm_cos_kernel(r);
    case 1: return -dm_sin_kernel(r);
    case 2: return -dm_cos_kernel(r);
    default: return dm_sin_kernel(r);
    }
}
DM_FN double dm_tan(double x) {
    int q;
    double r = dm_reduce_pio2(x, &q);
    double s = dm_sin_kernel(r), c = dm_cos_kernel(r);
    return (q & 1) ? -c / s : s / c;
}

DM_FN double dm_sqrt(double x);

DM_FN double dm_atan(double x) {
    double sgn = 1.0;
    if (x < 0.0) { x = -x; sgn = -1.0; }
    int inv = 0;
    if (x > 1.0) { x = 1.0 / x; inv = 1; }
    double scale = 1.0;
    int step;
    for (step = 0; step < 2 && x > 0.2; ++step) {
        x = x / (1.0 + dm_sqrt(1.0 + x * x));
        scale = scale * 2.0;
    }
    double x2 = x * x;
    double p = 1.0 / 29.0;
    p = -p * x2 + 1.0 / 27.0;
    p = -p * x2 + 1.0 / 25.0;
    p = -p * x2 + 1.0 / 23.0;
    p = -p * x2 + 1.0 / 21.0;
    p = -p * x2 + 1.0 / 19.0;
    p = -p * x2 + 1.0 / 17.0;
    p = -p * x2 + 1.0 / 15.0;
    p = -p * x2 + 1.0 / 13.0;
    p = -p * x2 + 1.0 / 11.0;
    p = -p * x2 + 1.0 / 9.0;
    p = -p * x2 + 1.0 / 7.0;
    p = -p * x2 + 1.0 / 5.0;
    p = -p * x2 + 1.0 / 3.0;
    p = -p * x2 + 1.0;
    double a = scale * (x * p);
    if (inv) a = DM_PIO2 - a;
    return sgn * a;
}

/* IEEE square root: both targets implement it correctly rounded. */
#if defined(__HIPCC__)
DM_FN double dm_sqrt(double x) { return __builtin_sqrt(x); }
DM_FN float dm_sqrtf(float x) { return __builtin_sqrtf(x); }
#else
#include <math.h>
DM_FN double dm_sqrt(double x) { return sqrt(x); }
DM_FN float dm_sqrtf(float x) { return sqrtf(x); }
#endif

/* ------------------------------------------------------------------
 * Single precision (the FP32 path): the same methods in float, with fewer
 * terms, built only from IEEE-754 float +, -, *, / and exact bit
 * manipulation, so the GPU and a CPU compiled with -ffp-contract=off agree
 * bit for bit.  Ranges and accuracy (against libm in double, rounded):
 *   dmf_log, dmf_log10 : x > 0 (0 gives -inf), about 1-2 ulp;
 *   dmf_exp            : any x (overflow to inf, underflow to 0), 1-2 ulp;
 *   dmf_pow(a, b)      : a > 0, exp(b log a), error grows with |b log a|
 *                        (a few ulp for |b log a| < 10);
 *   dmf_sincos         : |x| < 1e4 (phases 2 pi R of the collisions), 1-2 ulp
 *                        of the larger of sin, cos;
 *   dmf_tan            : |x| < 1e4, sin / cos;
 *   dmf_atan           : any x, 1-3 ulp.
 * Methods:
 *   log : x = m 2^e with m in [sqrt(1/2), sqrt(2)), s = (m-1)/(m+1),
 *         log m = 2 atanh(s) by its odd series to s^9; e ln 2 in two parts.
 *   exp : x = k ln 2 + r (two-part ln 2), |r| <= ln2/2, Taylor series of
 *         exp(r) to degree 7, scaling by 2^k through the exponent bits.
 *   sin/cos : reduction by pi/2 (three-part constant), Taylor series on
 *         |r| <= pi/4 (sin to r^9, cos to r^10), quadrant selection.
 *   atan : atan(x) = pi/2 - atan(1/x) for x > 1, atan(x) = pi/4 +
 *         atan((x-1)/(x+1)) for x > tan(pi/8), odd series to x^19.
 * ------------------------------------------------------------------ */
DM_FN uint32_t dmf_bits(float x) {
    union { float f; uint32_t u; } v;
    v.f = x;
    return v.u;
}
DM_FN float dmf_from_bits(uint32_t u) {
    union { float f; uint32_t u; } v;
    v.u = u;
    return v.f;
}

#define DMF_LN2_HI 0.693145751953125f        /* 0x3f317200: 15 significant bits */
#define DMF_LN2_LO 1.428606765330187e-06f    /* ln 2 - DMF_LN2_HI */
#define DMF_INV_LN2 1.44269502162933349609375f
#define DMF_INV_LN10 0.4342944920063019f
#define DMF_PIO2_1 1.5703125f                /* pi/2 in three parts (8 + 8 + 24 bits) */
#define DMF_PIO2_2 4.837512969970703125e-04f
#define DMF_PIO2_3 7.549789948768648e-08f
#define DMF_PIO2 1.57079637050628662109375f
#define DMF_PIO4 0.785398185253143310546875f

DM_FN float dmf_log(float x) {
    if (!(x > 0.0f)) return x == 0.0f ? -dmf_from_bits(0x7f800000u) : dmf_from_bits(0x7fc00000u);
    uint32_t u = dmf_bits(x);
    if (u >= 0x7f800000u) return x;   /* inf, nan */
    int e = (int)(u >> 23) - 127;
    if (e == -127) {                  /* subnormal: scale by 2^24 */
        u = dmf_bits(x * 16777216.0f);
        e = (int)(u >> 23) - 127 - 24;
    }
    float m = dmf_from_bits((u & 0x007fffffu) | 0x3f800000u);   /* [1, 2) */
    if (m > 1.41421353816986083984375f) {
        m = m * 0.5f;
        e = e + 1;
    }
    float s = (m - 1.0f) / (m + 1.0f);
    float z = s * s;
    float p = 0.111111111f;           /* 1/9 */
    p = p * z + 0.142857143f;         /* 1/7 */
    p = p * z + 0.2f;
    p = p * z + 0.333333333f;
    float lm = (s + s) + (s + s) * (z * p);
    float fe = (float)e;
    return fe * DMF_LN2_HI + (lm + fe * DMF_LN2_LO);
}

DM_FN float dmf_log10(float x) { return dmf_log(x) * DMF_INV_LN10; }

DM_FN float dmf_exp(float x) {
    if (x != x) return x;
    if (x > 88.7228f) return dmf_from_bits(0x7f800000u);
    if (x < -103.98f) return 0.0f;
    float t = x * DMF_INV_LN2;
    int k = (int)(t < 0.0f ? t - 0.5f : t + 0.5f);
    float fk = (float)k;
    float r = (x - fk * DMF_LN2_HI) - fk * DMF_LN2_LO;
    float p = 1.98412698e-04f;        /* 1/5040 */
    p = p * r + 1.38888889e-03f;      /* 1/720 */
    p = p * r + 8.33333333e-03f;      /* 1/120 */
    p = p * r + 4.16666667e-02f;      /* 1/24 */
    p = p * r + 0.166666667f;         /* 1/6 */
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    /* scale by 2^k: two steps keep every factor a normal power of two */
    if (k < -125) {
        p = p * dmf_from_bits((uint32_t)(k + 100 + 127) << 23);
        return p * dmf_from_bits((uint32_t)(-100 + 127) << 23);
    }
    if (k > 127) {
        p = p * dmf_from_bits((uint32_t)(k - 1 + 127) << 23);
        return p * 2.0f;
    }
    return p * dmf_from_bits((uint32_t)(k + 127) << 23);
}

DM_FN float dmf_pow(float a, float b) { return dmf_exp(b * dmf_log(a)); }

/* x = q pi/2 + r, |r| <= pi/4 (for |x| < 1e4) */
DM_FN float dmf_reduce_pio2(float x, int *q) {
    float t = x * 0.636619772f;       /* 2/pi */
    int k = (int)(t < 0.0f ? t - 0.5f : t + 0.5f);
    float fk = (float)k;
    *q = k & 3;
    return ((x - fk * DMF_PIO2_1) - fk * DMF_PIO2_2) - fk * DMF_PIO2_3;
}
DM_FN float dmf_sin_kernel(float r) {
    float z = r * r;
    float p = 2.75573192e-06f;        /* 1/9! */
    p = -p * z + 1.98412698e-04f;     /* 1/7! */
    p = -p * z + 8.33333333e-03f;     /* 1/5! */
    p = -p * z + 0.166666667f;        /* 1/3! */
    return r - r * (z * p);
}
DM_FN float dmf_cos_kernel(float r) {
    float z = r * r;
    float p = 2.75573192e-07f;        /* 1/10! */
    p = -p * z + 2.48015873e-05f;     /* 1/8! */
    p = -p * z + 1.38888889e-03f;     /* 1/6! */
    p = -p * z + 4.16666667e-02f;     /* 1/4! */
    p = -p * z + 0.5f;
    return 1.0f - z * p;
}
DM_FN void dmf_sincos(float x, float *sn, float *cs) {
    int q;
    float r = dmf_reduce_pio2(x, &q);
    float s = dmf_sin_kernel(r), c = dmf_cos_kernel(r);
    switch (q) {
    case 0: *sn = s; *cs = c; break;
    case 1: *sn = c; *cs = -s; break;
    case 2: *sn = -s; *cs = -c; break;
    default: *sn = -c; *cs = s; break;
    }
}
DM_FN float dmf_sin(float x) { float s, c; dmf_sincos(x, &s, &c); return s; }
DM_FN float dmf_cos(float x) { float s, c; dmf_sincos(x, &s, &c); return c; }
DM_FN float dmf_tan(float x) {
    int q;
    float r = dmf_reduce_pio2(x, &q);
    float s = dmf_sin_kernel(r), c = dmf_cos_kernel(r);
    return (q & 1) ? -c / s : s / c;
}

DM_FN float dmf_atan(float x) {
    float sgn = 1.0f;
    if (x < 0.0f) { x = -x; sgn = -1.0f; }
    if (x != x) return x;
    int inv = 0;
    if (x > 1.0f) { x = 1.0f / x; inv = 1; }
    int shift = 0;
    if (x > 0.414213562f) { x = (x - 1.0f) / (x + 1.0f); shift = 1; }
    float z = x * x;
    float p = 0.0526315789f;          /* 1/19 */
    p = -p * z + 0.0588235294f;       /* 1/17 */
    p = -p * z + 0.0666666667f;       /* 1/15 */
    p = -p * z + 0.0769230769f;       /* 1/13 */
    p = -p * z + 0.0909090909f;       /* 1/11 */
    p = -p * z + 0.111111111f;        /* 1/9 */
    p = -p * z + 0.142857143f;        /* 1/7 */
    p = -p * z + 0.2f;
    p = -p * z + 0.333333333f;
    float a = x - x * (z * p);
    if (shift) a = DMF_PIO4 + a;
    if (inv) a = DMF_PIO2 - a;
    return sgn * a;
}

#endif /* PIC_DETMATH_H */
