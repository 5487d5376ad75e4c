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
 *  - fm_sincos: Cody-Waite reduction by pi/2 with a three-part constant (the
 *    first part subtracted exactly by one fma), carrying the reduced argument
 *    as a double-double; minimax polynomials of sin and cos on [-pi/4, pi/4]
 *    (degree 13 and 14, six coefficients each, relative error 2^-57.9 and
 *    2^-64.0, Remez exchange in gen_fm_consts.py) evaluated by fused
 *    multiply-adds, the cosine assembled with the exact correction of
 *    1 - x^2/2 (the technique of the fdlibm kernels).
 *  - fm_sincos_near / fm_sincos3: sin and cos at points within 2^-12 of x
 *    (the theta stencil of the finite differences) from x's reduced sine and
 *    cosine carried as double-doubles, by the addition formulas with short
 *    Taylor polynomials of the offset.
 *  - fm_exp: P. T. P. Tang, "Table-driven implementation of the exponential
 *    function in IEEE floating-point arithmetic", ACM TOMS 15 (1989): x =
 *    (64k + j) ln2/64 + r, exp(x) = 2^k 2^(j/64) exp(r), |r| <= ln2/128,
 *    2^(j/64) stored as hi + lo, exp(r) - 1 by its Taylor polynomial to r^6.
 *  - fm_pow2/3/4/8: integer powers by double-double products (Dekker's exact
 *    product via fma), so the result is the correctly rounded power except in
 *    rare near-midpoint cases.
 * Accuracy against mpmath (tests/test_fm_math.py): sin and cos within 0.69 ulp
 * on the path's arguments (about 1.5% of results not correctly rounded), exp
 * within 0.55 ulp, the integer powers correctly rounded but for rare
 * near-midpoint cases.
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

/* A polynomial coefficient as an operand of a fused multiply-add.  On the GPU the value is
   placed in a scalar register pair, which the multiply-add reads directly; as a literal the
   compiler copies it into vector registers first (two extra vector instructions per term).
   The empty statement emits no instruction and leaves the value unchanged. */
#if defined(__HIP_DEVICE_COMPILE__)
FM_FN double fm_k(double c)
{
    __asm__("" : "+s"(c));
    return c;
}
#else
#define fm_k(c) (c)
#endif

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

/* The reduced sine and cosine of x as unevaluated sums: x = n pi/2 + t with n the nearest
   integer to x 2/pi (returned mod 4), sin t = sh + sl and cos t = ch + cl, each low part the
   exact rounding residual of its high part.  Arguments |x| < 2^30 are reduced with full
   accuracy. */
FM_FN int fm_sincos_red(double x, double *sh, double *sl, double *ch, double *cl)
{
    if (!(x - x == 0.0)) { /* NaN or infinity */
        *sh = *sl = *ch = *cl = x - x;
        return 0;
    }
    /* x = kd pi/2 + (hi + lo).  kd is the integer nearest x 2/pi.  x - kd P1 is exact in one
       fused multiply-add: for |x| < 2^30, x and kd P1 are both multiples of 2^-53 (|x| < 1)
       or of 2^-52 (|x| >= 1) and differ by less than 0.8, so the difference fits 53 bits.
       The two lower parts of pi/2 follow as a double-double. */
    double kd = __builtin_fma(x, FM_TWO_OVER_PI, FM_SHIFT) - FM_SHIFT;
    double r = __builtin_fma(-kd, FM_PIO2_1, x);
    double cc = __builtin_fma(kd, FM_PIO2_2, kd * FM_PIO2_3);
    double lo;
    double hi = fm_two_sum(r, -cc, &lo);

    double z = hi * hi;
    double hz = 0.5 * z;
    /* sin(hi + lo) = hi + hi^3 S(z) + lo (1 - z/2) */
    double ps = __builtin_fma(z, FM_S13, fm_k(FM_S11));
    ps = __builtin_fma(z, ps, fm_k(FM_S9));
    ps = __builtin_fma(z, ps, fm_k(FM_S7));
    ps = __builtin_fma(z, ps, fm_k(FM_S5));
    ps = __builtin_fma(z, ps, fm_k(FM_S3));
    double ts = __builtin_fma(hi * z, ps, __builtin_fma(-hz, lo, lo));
    double sn = hi + ts;
    /* cos(hi + lo) = 1 - z/2 + z^2 C(z) - hi lo, with hi^2 = z + zl exactly; 1 - z/2 as
       w + ((1 - w) - z/2) with the exact correction of w = 1 - z/2 (the fdlibm kernels' device) */
    double zl = __builtin_fma(hi, hi, -z);
    double pc = __builtin_fma(z, FM_C14, fm_k(FM_C12));
    pc = __builtin_fma(z, pc, fm_k(FM_C10));
    pc = __builtin_fma(z, pc, fm_k(FM_C8));
    pc = __builtin_fma(z, pc, fm_k(FM_C6));
    pc = __builtin_fma(z, pc, fm_k(FM_C4));
    pc = (z * z) * pc;
    double w = 1.0 - hz;
    double tc = __builtin_fma(-0.5, zl, (1.0 - w) - hz) + __builtin_fma(-hi, lo, pc);
    double cs = w + tc;
    /* the residuals by Fast2Sum, exact since |ts| < |hi| (|ts| <= 0.11 |hi| + |lo|, |lo| <=
       ulp(hi)/2) and |tc| < 0.02 < |w| */
    *sh = sn;
    *sl = ts - (sn - hi);
    *ch = cs;
    *cl = tc - (cs - w);
    return ((int)kd) & 3;
}

/* sin and cos of n pi/2 + t from sn = sin t, cs = cos t: (sn, cs), (cs, -sn), (-sn, -cs),
   (-cs, sn) for n = 0..3, selected without branches (only exact moves and negations) */
FM_FN void fm_quadrant(int n, double sn, double cs, double *s, double *c)
{
    int odd = n & 1;
    double a = odd ? cs : sn;
    double b = odd ? sn : cs;
    *s = (n & 2) ? -a : a;
    *c = ((n + 1) & 2) ? -b : b;
}

/* sin(x) and cos(x); arguments |x| < 2^30 are reduced with full accuracy.  The operations of
   fm_sincos_red and fm_quadrant, written out (the same values; a NaN or infinity returns
   before the quadrant selection, which keeps the kernels' code as it was before the reduced
   form was split out). */
FM_FN void fm_sincos(double x, double *s, double *c)
{
    if (!(x - x == 0.0)) { /* NaN or infinity */
        *s = x - x;
        *c = x - x;
        return;
    }
    double kd = __builtin_fma(x, FM_TWO_OVER_PI, FM_SHIFT) - FM_SHIFT;
    double r = __builtin_fma(-kd, FM_PIO2_1, x);
    double cc = __builtin_fma(kd, FM_PIO2_2, kd * FM_PIO2_3);
    double lo;
    double hi = fm_two_sum(r, -cc, &lo);

    double z = hi * hi;
    double hz = 0.5 * z;
    double ps = __builtin_fma(z, FM_S13, fm_k(FM_S11));
    ps = __builtin_fma(z, ps, fm_k(FM_S9));
    ps = __builtin_fma(z, ps, fm_k(FM_S7));
    ps = __builtin_fma(z, ps, fm_k(FM_S5));
    ps = __builtin_fma(z, ps, fm_k(FM_S3));
    double sn = hi + __builtin_fma(hi * z, ps, __builtin_fma(-hz, lo, lo));
    double zl = __builtin_fma(hi, hi, -z);
    double pc = __builtin_fma(z, FM_C14, fm_k(FM_C12));
    pc = __builtin_fma(z, pc, fm_k(FM_C10));
    pc = __builtin_fma(z, pc, fm_k(FM_C8));
    pc = __builtin_fma(z, pc, fm_k(FM_C6));
    pc = __builtin_fma(z, pc, fm_k(FM_C4));
    pc = (z * z) * pc;
    double w = 1.0 - hz;
    double cs = w + (__builtin_fma(-0.5, zl, (1.0 - w) - hz) + __builtin_fma(-hi, lo, pc));

    int n = ((int)kd) & 3;
    int odd = n & 1;
    double a = odd ? cs : sn;
    double b = odd ? sn : cs;
    *s = (n & 2) ? -a : a;
    *c = ((n + 1) & 2) ? -b : b;
}

/* Whether fm_sincos_near shifts y from x: |d| <= 2^-12 for d = y - x, and |x| >= 2^-11, so
   that d is exact (Sterbenz's lemma) */
FM_FN int fm_near_ok(double x, double y)
{
    return __builtin_fabs(y - x) <= 0x1p-12 && __builtin_fabs(x) >= 0x1p-11;
}

/* sin(t + d) and cos(t + d) from the reduced sine and cosine of t (sh + sl, ch + cl) by the
   addition formulas, mapped to quadrant n:
     sin(t + d) = sh + (sl + sh (cos d - 1) + ch sin d),
     cos(t + d) = ch + (cl + ch (cos d - 1) - sh sin d),
   the bracket (of the order of d) evaluated by fused multiply-adds and added once, with
   sin d = d - d^3/6 + d^5/120 and cos d - 1 = -d^2/2 + d^4/24 (truncation below 2^-80
   relative for |d| <= 2^-12) and the products sl (cos d - 1) and cl sin d dropped (below
   2^-64 |d| in absolute terms). */
FM_FN void fm_sincos_shift(int n, double sh, double sl, double ch, double cl, double d, double *s,
                           double *c)
{
    double z = d * d;
    double sd = __builtin_fma(d * z, __builtin_fma(z, 1.0 / 120.0, -1.0 / 6.0), d);
    double cm1 = z * __builtin_fma(z, 1.0 / 24.0, -0.5);
    double sn = sh + __builtin_fma(ch, sd, __builtin_fma(sh, cm1, sl));
    double cs = ch + __builtin_fma(-sh, sd, __builtin_fma(ch, cm1, cl));
    fm_quadrant(n, sn, cs, s, c);
}

/* sin(y) and cos(y) for y near x, from the reduced sine and cosine of x (fm_sincos_red:
   quadrant n, sh + sl, ch + cl): shifted by d = y - x (fm_sincos_shift) where fm_near_ok,
   evaluated in full elsewhere.  The result differs from fm_sincos(y) only by the rounding of
   the bracket and the accuracy of the polynomials behind the double-double (about 1.6% of
   results differ in the last bit, as accurate: within 0.75 ulp, with an absolute 2^-52 |d|
   added where the result is close to zero). */
FM_FN void fm_sincos_near(int n, double sh, double sl, double ch, double cl, double x, double y,
                          double *s, double *c)
{
    if (fm_near_ok(x, y))
        fm_sincos_shift(n, sh, sl, ch, cl, y - x, s, c);
    else
        fm_sincos(y, s, c);
}

/* sin and cos at x and at two points xp, xm near it (the finite-difference stencil in theta):
   each of xp and xm as fm_sincos_near gives it, with one reduction and one pair of polynomials
   for the three where both shift (the full evaluations in one rarely taken branch) */
FM_FN void fm_sincos3(double x, double xp, double xm, double *s0, double *c0, double *sp, double *cp,
                      double *sm, double *cm)
{
    double sh, sl, ch, cl;
    int n = fm_sincos_red(x, &sh, &sl, &ch, &cl);
    fm_quadrant(n, sh, ch, s0, c0);
    fm_sincos_shift(n, sh, sl, ch, cl, xp - x, sp, cp);
    fm_sincos_shift(n, sh, sl, ch, cl, xm - x, sm, cm);
    int okp = fm_near_ok(x, xp), okm = fm_near_ok(x, xm);
    if (!(okp && okm)) {
        double s1, c1, s2, c2;
        fm_sincos(xp, &s1, &c1);
        fm_sincos(xm, &s2, &c2);
        if (!okp) {
            *sp = s1;
            *cp = c1;
        }
        if (!okm) {
            *sm = s2;
            *cm = c2;
        }
    }
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
