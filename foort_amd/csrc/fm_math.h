/*
 * fm_math.h -- portable IEEE-754 binary64 elementary functions for the
 * geodesic integration path.
 *
 * The reference evaluates its metrics with the C library's sin, cos, exp and
 * pow (FOORT/Metric.cpp:134-191, :276-361, :409-473, :511-626) and its
 * finite-difference Christoffel symbols amplify the last bit of every metric
 * value by 1/(2h) = 5e6 (FOORT/Metric.cpp:30).  GPU and CPU libraries round
 * differently in the last bit, so the same trajectory computed with two
 * libraries drifts apart.  This header is the single definition of those
 * functions for both sides: it is plain C99 that compiles unchanged with gcc
 * for x86-64 and with hipcc for gfx950, uses only correctly rounded IEEE
 * operations (+ - * / sqrt, fused multiply-add where written explicitly) and
 * must be compiled with floating-point contraction off.  Under those rules
 * every function returns bit-identical results on host and device.
 *
 * Methods (all constants come from gen_fm_consts.py, computed with mpmath):
 *  - fm_sincos: Cody-Waite reduction by pi/2 with a three-part constant and an
 *    exact product split by fma, carrying the reduced argument as a
 *    double-double; Taylor polynomials of sin and cos on [-pi/4, pi/4]
 *    (degree 17 and 18, truncation error < 2^-62), the cosine assembled with
 *    the exact correction of 1 - x^2/2 (the technique of the fdlibm kernels).
 *  - fm_exp: P. T. P. Tang, "Table-driven implementation of the exponential
 *    function in IEEE floating-point arithmetic", ACM TOMS 15 (1989): x =
 *    (64k + j) ln2/64 + r, exp(x) = 2^k 2^(j/64) exp(r), |r| <= ln2/128,
 *    2^(j/64) stored as hi + lo, exp(r) - 1 by its Taylor polynomial to r^6.
 *  - fm_pow2/3/4/8: integer powers by double-double products (Dekker's exact
 *    product via fma), so the result is the correctly rounded power except in
 *    rare near-midpoint cases.
 * Accuracy (tests/test_fm_math.py, against mpmath): below 0.52 ulp on the
 * argument ranges the path uses.
 */
#ifndef FOORT_FM_MATH_H
#define FOORT_FM_MATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#define FM_FN __host__ __device__ static inline
#else
#define FM_FN static inline
#endif

#include "fm_consts.inc"

/* round-to-nearest-even of |y| < 2^51 by the 1.5*2^52 shift */
#define FM_SHIFT 0x1.8p52

typedef union {
    double d;
    uint64_t u;
} fm_bits;

FM_FN double fm_from_bits(uint64_t u)
{
    fm_bits b;
    b.u = u;
    return b.d;
}

/* 2^k for -1022 <= k <= 1023 */
FM_FN double fm_pow2i(int k)
{
    return fm_from_bits((uint64_t)(k + 1023) << 52);
}

/* Knuth's TwoSum: s + e == a + b exactly */
FM_FN double fm_two_sum(double a, double b, double *e)
{
    double s = a + b;
    double bb = s - a;
    *e = (a - (s - bb)) + (b - bb);
    return s;
}

/* exp(x) */
FM_FN double fm_exp(double x)
{
    static const double T_HI[64] = FM_EXP2_TABLE_HI;
    static const double T_LO[64] = FM_EXP2_TABLE_LO;
    if (!(x < 0x1.62e42fefa39efp+9)) {   /* x >= ln(DBL_MAX) (709.78...) or NaN */
        if (x != x)
            return x + x;
        return 0x1.fffffffffffffp+1023 * 2.0; /* +inf */
    }
    if (x < -0x1.74910d52d3052p+9) /* x < -745.13...: below half the least subnormal */
        return 0.0;
    double md = (x * FM_INV_LN2_64 + FM_SHIFT) - FM_SHIFT;
    int m = (int)md;
    int j = m & 63;
    int k = (m - j) / 64;
    double t = md * FM_LN2_64_1;
    double e = __builtin_fma(md, FM_LN2_64_1, -t); /* md*c1 == t + e exactly */
    double r = x - t;                              /* exact (Sterbenz) */
    double c = e + md * FM_LN2_64_2;
    double lo;
    double hi = fm_two_sum(r, -c, &lo);
    double q = hi + hi * hi * (FM_E2 + hi * (FM_E3 + hi * (FM_E4 + hi * (FM_E5 + hi * FM_E6))));
    q = q + lo;
    double y = T_HI[j] + (T_HI[j] * q + T_LO[j]);
    /* scale by 2^k in two exact steps so only the final product rounds */
    int k1 = k / 2;
    int k2 = k - k1;
    return (y * fm_pow2i(k1)) * fm_pow2i(k2);
}

/* sin(x) and cos(x); arguments |x| < 2^30 are reduced with full accuracy */
FM_FN void fm_sincos(double x, double *s, double *c)
{
    if (!(x - x == 0.0)) { /* NaN or infinity */
        *s = x - x;
        *c = x - x;
        return;
    }
    double kd = (x * FM_TWO_OVER_PI + FM_SHIFT) - FM_SHIFT;
    double t = kd * FM_PIO2_1;
    double e = __builtin_fma(kd, FM_PIO2_1, -t); /* kd*P1 == t + e exactly */
    double r = x - t;                             /* exact (Sterbenz) */
    double cc = e + kd * FM_PIO2_2;
    cc = cc + kd * FM_PIO2_3;
    double lo;
    double hi = fm_two_sum(r, -cc, &lo);

    double z = hi * hi;
    /* sin(hi + lo) = hi + hi^3 S(z) + lo (1 - z/2) */
    double ps = FM_S5 + z * (FM_S7 + z * (FM_S9 + z * (FM_S11 + z * (FM_S13 + z * (FM_S15 + z * FM_S17)))));
    double sn = hi + ((hi * z) * (FM_S3 + z * ps) + lo * (1.0 - 0.5 * z));
    /* cos(hi + lo) = 1 - z/2 + z^2 C(z) - hi lo, with z = zh + zl exactly */
    double zl = __builtin_fma(hi, hi, -z);
    double pc = (z * z) * (FM_C4 + z * (FM_C6 + z * (FM_C8 + z * (FM_C10 + z * (FM_C12 + z * (FM_C14 + z * (FM_C16 + z * FM_C18)))))));
    double hz = 0.5 * z;
    double w = 1.0 - hz;
    double cs = w + ((((1.0 - w) - hz) - 0.5 * zl) + (pc - hi * lo));

    /* quadrant n: (s, c) = (sn, cs), (cs, -sn), (-sn, -cs), (-cs, sn) for n = 0..3;
       selected without branches (only exact moves and negations) */
    int n = ((int)kd) & 3;
    int odd = n & 1;
    double a = odd ? cs : sn;
    double b = odd ? sn : cs;
    *s = (n & 2) ? -a : a;
    *c = ((n + 1) & 2) ? -b : b;
}

FM_FN double fm_sin(double x)
{
    double s, c;
    fm_sincos(x, &s, &c);
    return s;
}

FM_FN double fm_cos(double x)
{
    double s, c;
    fm_sincos(x, &s, &c);
    return c;
}

/* x^2: a single correctly rounded product */
FM_FN double fm_pow2(double x)
{
    return x * x;
}

/* x^3 via x^2 = h + l exactly, then (h + l) x in double-double */
FM_FN double fm_pow3(double x)
{
    double h = x * x;
    double l = __builtin_fma(x, x, -h);
    double t = h * x;
    double tl = __builtin_fma(h, x, -t);
    return t + (tl + l * x);
}

/* x^4 = (h + l)^2 in double-double; *lo receives the low part */
FM_FN double fm_pow4_dd(double x, double *lo)
{
    double h = x * x;
    double l = __builtin_fma(x, x, -h);
    double t = h * h;
    double tl = __builtin_fma(h, h, -t);
    double a = tl + (2.0 * h) * l;
    double s = t + a;
    *lo = a - (s - t);
    return s;
}

FM_FN double fm_pow4(double x)
{
    double lo;
    return fm_pow4_dd(x, &lo);
}

/* x^8 = (x^4)^2 with x^4 carried as a double-double */
FM_FN double fm_pow8(double x)
{
    double a;
    double A = fm_pow4_dd(x, &a);
    double B = A * A;
    double bl = __builtin_fma(A, A, -B);
    return B + (bl + (2.0 * A) * a);
}

#endif /* FOORT_FM_MATH_H */
