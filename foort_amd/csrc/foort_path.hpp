// foort_path.hpp -- the per-pixel geodesic path, compiled for gfx950 kernels
// and (for the camera set-up) for the host.
//
// Everything here follows the reference's arithmetic term by term so that,
// compiled with floating-point contraction off, the GPU reproduces the CPU
// oracle bit for bit:
//   metrics            FOORT/Metric.cpp:131-191, :218-229, :273-361, :409-473, :511-626
//   Christoffel (FD)   FOORT/Metric.cpp:14-55   (sparse form: contraction.inc)
//   adaptive step      FOORT/Integrators.cpp:34-47
//   RK4 / Verlet       FOORT/Integrators.cpp:51-94, :99-152
//   Geodesic::Update   FOORT/Geodesic.cpp:61-88
//   terminations       FOORT/Terminations.cpp:71-214
//   diagnostics        FOORT/Diagnostics.cpp:113-140, :159-200, :351-380, :427-446
//   initial conditions FOORT/ViewScreen.cpp:75-119
// The metric expressions are written exactly as the reference writes them
// (same operands, same association) because every rounding matters.
#pragma once

#include <stdint.h>

#include "dual.hpp"
#include "fm_div.hpp"
#include "fm_math.h"
#include "foort_gpu.h"

#if defined(__HIPCC__)
#define FM_HD __host__ __device__
#define FM_HD_INLINE __host__ __device__ __attribute__((always_inline)) inline
#else
#define FM_HD
#define FM_HD_INLINE inline
#endif

namespace foort {

constexpr double kPi = 3.1415926535; // FOORT/Geometry.h:44

// ---------------------------------------------------------------------------
// Elementary functions for the scalar types of the path.  double is the
// reference-exact type (the portable fm_math.h functions); D2<S> carries
// derivatives (analytic Christoffel symbols), S = double or float.
// ---------------------------------------------------------------------------
FM_HD_INLINE double fexp(double x) { return fm_exp(x); }
FM_HD_INLINE double fsqrt(double x) { return __builtin_sqrt(x); }
FM_HD_INLINE double fpow2(double x) { return fm_pow2(x); }
FM_HD_INLINE double fpow3(double x) { return fm_pow3(x); }
FM_HD_INLINE double fpow4(double x) { return fm_pow4(x); }
FM_HD_INLINE double fpow8(double x) { return fm_pow8(x); }

FM_HD_INLINE float sexp(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __expf(x);
#else
    return __builtin_expf(x);
#endif
}
// binary64 exponential of the analytic formulations (not the reference's arithmetic, so not
// bound to fm_exp's bits): Cody and Waite's reduction x = k ln2 + r, |r| <= ln2/2, with ln2 in
// two parts applied by fused multiply-adds, then exp(r) by its Taylor polynomial to r^13
// (truncation below 5e-18 relative) and the scaling by 2^k with ldexp.  No table, so no
// memory load: a lone wave does not wait on one.  Error within 2 ulp (tested against
// numpy's exp over the path's argument range).
FM_HD_INLINE double sexp(double x)
{
    if (!(x < 709.782712893384)) // overflow, +inf or NaN
        return x + __builtin_huge_val();
    if (x < -745.2)
        return 0.0;
    const double k = __builtin_rint(x * 1.4426950408889634);
    double r = __builtin_fma(-k, 0x1.62e42fefa39efp-1, x);
    r = __builtin_fma(-k, 0x1.abc9e3b39803fp-56, r);
    double p = 1.0 / 6227020800.0;            // 1/13!
    p = __builtin_fma(p, r, fm_k(1.0 / 479001600.0)); // 1/12!
    p = __builtin_fma(p, r, fm_k(1.0 / 39916800.0));
    p = __builtin_fma(p, r, fm_k(1.0 / 3628800.0));
    p = __builtin_fma(p, r, fm_k(1.0 / 362880.0));
    p = __builtin_fma(p, r, fm_k(1.0 / 40320.0));
    p = __builtin_fma(p, r, fm_k(1.0 / 5040.0));
    p = __builtin_fma(p, r, fm_k(1.0 / 720.0));
    p = __builtin_fma(p, r, fm_k(1.0 / 120.0));
    p = __builtin_fma(p, r, fm_k(1.0 / 24.0));
    p = __builtin_fma(p, r, fm_k(1.0 / 6.0));
    p = __builtin_fma(p, r, 0.5);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    return __builtin_ldexp(p, static_cast<int>(k));
}
FM_HD_INLINE float ssqrt(float x) { return __builtin_sqrtf(x); }
FM_HD_INLINE double ssqrt(double x) { return __builtin_sqrt(x); }
FM_HD_INLINE void ssincos(double x, double *s, double *c) { fm_sincos(x, s, c); }
FM_HD_INLINE void ssincos(float x, float *s, float *c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    __sincosf(x, s, c);
#else
    *s = __builtin_sinf(x);
    *c = __builtin_cosf(x);
#endif
}
FM_HD_INLINE void ssincos(double x, float *s, float *c) { ssincos(static_cast<float>(x), s, c); }

template <class S>
FM_HD_INLINE D2<S> fexp(const D2<S> &a)
{
    const S e = sexp(a.v);
    return D2<S>(e, e * a.du, e * a.dt);
}
template <class S>
FM_HD_INLINE D2<S> fsqrt(const D2<S> &a)
{
    const S r = ssqrt(a.v);
    const S k = S(0.5) * arcp(r);
    return D2<S>(r, a.du * k, a.dt * k);
}
template <class S>
FM_HD_INLINE D2<S> fpow2(const D2<S> &a) { return a * a; }
template <class S>
FM_HD_INLINE D2<S> fpow3(const D2<S> &a) { return a * a * a; }
template <class S>
FM_HD_INLINE D2<S> fpow4(const D2<S> &a)
{
    const D2<S> b = a * a;
    return b * b;
}
template <class S>
FM_HD_INLINE D2<S> fpow8(const D2<S> &a)
{
    const D2<S> b = fpow4(a);
    return b * b;
}

// ---------------------------------------------------------------------------
// Run constants (filled on the host by foort_gpu_open, passed by value)
// ---------------------------------------------------------------------------
struct MetricParams {
    double a;               // Kerr / Johannsen / MN spin, RL rescaled a
    double m, p, q;         // Rasheed-Larsen (rescaled by (p+q)/4)
    double rl_root_pq;      // sqrt(p q)                        (computed once on the host, where no
    double rl_root_pm;      // sqrt((p^2 - 4m^2) (q^2 - 4m^2))   multiply-add can make p = 2m negative)
    double a13, a22, a52, eps3;  // Johannsen
    double alpha3, mn_alpha, mn_k;  // Manko-Novikov
    double horizon;         // SphericalHorizonMetric::m_HorizonRadius
    int32_t rlog;
    int32_t has_horizon;
};

struct RunConst {
    MetricParams M;
    // integrator
    double eps, minstep, h_fd, inv2h, verlet_tol;
    // terminations
    uint32_t terms;
    uint32_t diags;
    double h_eps, sphere_R, theta_eps;
    uint32_t h_n, s_n, t_n, th_n;   // update every n steps (0 = always)
    uint32_t max_steps;
    uint32_t step_cap;
    // diagnostics
    double ep_thr;
    uint32_t ep_n, cr_n;
    int32_t ep_start, ep_finish, cr_start, cr_finish;
    // camera
    double V[4][4];        // vielbein, row a = e_a
    double g0[4];          // g_dd[0][i] at the camera
    double cam_r;          // true camera radius
    double cam_u;          // radial coordinate of the camera (log r with RLogScale)
    double cam_t, cam_theta, cam_phi;
    double sin_theta0;     // sin(camera theta)
    double size_x, size_y, center_x, center_y;
    // mixed precision: binary32 right-hand side where u > mixed_u and theta is
    // farther than mixed_pole from both poles
    double mixed_u, mixed_pole;
};

// ---------------------------------------------------------------------------
// Metrics: g_dd and g^uu at (u, theta) given sin/cos of theta.  u is the
// radial coordinate (r, or log r with RLogScale).  Output order g00 g03 g11 g22 g33.
// ---------------------------------------------------------------------------
FM_HD_INLINE double true_r(const MetricParams &M, double u)
{
    return M.rlog ? fm_exp(u) : u;
}

// FOORT/Metric.cpp:131-159
template <class T>
FM_HD_INLINE void kerr_dd(const MetricParams &M, T u, T sint, T cost, T *g)
{
    const double a = M.a;
    T r = M.rlog ? fexp(u) : u;
    T sigma = r * r + a * a * cost * cost;
    T delta = r * r + a * a - 2. * r;
    T A_ = (r * r + a * a) * (r * r + a * a) - delta * a * a * sint * sint;
    T g00 = -(1. - fm_div(2. * r, sigma));
    T g11 = fm_div(sigma, delta);
    T g22 = sigma;
    T g33 = fm_div(A_, sigma) * sint * sint;
    T g03 = fm_div(-2. * a * r * sint * sint, sigma);
    if (M.rlog)
        g11 *= (r * r);
    g[0] = g00; g[1] = g03; g[2] = g11; g[3] = g22; g[4] = g33;
}

// FOORT/Metric.cpp:162-191
template <class T>
FM_HD_INLINE void kerr_uu(const MetricParams &M, T u, T sint, T cost, T *g)
{
    const double a = M.a;
    T r = M.rlog ? fexp(u) : u;
    T sigma = r * r + a * a * cost * cost;
    T delta = r * r + a * a - 2. * r;
    T A_ = (r * r + a * a) * (r * r + a * a) - delta * a * a * sint * sint;
    T g00 = fm_div(-A_, sigma * delta);
    T g11 = fm_div(delta, sigma);
    T g22 = fm_div(1., sigma);
    T g33 = fm_div(delta - a * a * sint * sint, sigma * delta * sint * sint);
    T g03 = fm_div(-2. * a * r, sigma * delta);
    if (M.rlog)
        g11 *= fm_div(1.0, r * r);
    g[0] = g00; g[1] = g03; g[2] = g11; g[3] = g22; g[4] = g33;
}

// FOORT/Metric.cpp:218-229 (no RLogScale: FlatSpace is not a SphericalHorizonMetric)
template <class T>
FM_HD_INLINE void flat_dd(const MetricParams &, T u, T sint, T, T *g)
{
    g[0] = -1; g[1] = 0; g[2] = 1; g[3] = u * u; g[4] = u * u * sint * sint;
}

template <class T>
FM_HD_INLINE void flat_uu(const MetricParams &, T u, T sint, T, T *g)
{
    g[0] = -1; g[1] = 0; g[2] = 1; g[3] = fm_div(1., u * u); g[4] = fm_div(1., u * u * sint * sint);
}

// FOORT/Metric.cpp:276-300 / :321-345
template <class T>
struct RLCommon {
    T r, sint2, delta, H3, Bp, H1, H2, sqH1H2;
};

template <class T>
FM_HD_INLINE void rl_common(const MetricParams &M, T u, T sint, T cost, RLCommon<T> &c)
{
    const double m = M.m, a = M.a, pp = M.p, qq = M.q;
    T r = M.rlog ? fexp(u) : u;
    T r2 = r * r;
    T sint2 = sint * sint;
    T cost2 = cost * cost;
    T delta = r2 + a * a - 2. * r * m;
    T H3 = r2 - 2. * r * m + a * a * cost2;
    T Bp = M.rl_root_pq * a * sint2 *
                fm_div((pp * qq + 4. * m * m) * r - m * (pp - 2. * m) * (qq - 2. * m), 2. * m * (pp + qq) * H3);
    T H1 = r2 + a * a * cost2 + r * (pp - 2. * m) + fm_div(pp, pp + qq) * ((pp - 2. * m) * (qq - 2. * m) / 2.) -
                fm_div(pp, 2. * m * (pp + qq)) * M.rl_root_pm * a * cost;
    T H2 = r2 + a * a * cost2 + r * (qq - 2. * m) + fm_div(qq, pp + qq) * ((pp - 2. * m) * (qq - 2. * m) / 2.) +
                fm_div(qq, 2. * m * (pp + qq)) * M.rl_root_pm * a * cost;
    c.r = r; c.sint2 = sint2; c.delta = delta; c.H3 = H3; c.Bp = Bp; c.H1 = H1; c.H2 = H2;
    c.sqH1H2 = fsqrt(H1 * H2);
}

template <class T>
FM_HD_INLINE void rl_dd(const MetricParams &M, T u, T sint, T cost, T *g)
{
    RLCommon<T> c;
    rl_common(M, u, sint, cost, c);
    T g00 = fm_div(-c.H3, c.sqH1H2);
    T g11 = fm_div(c.sqH1H2, c.delta);
    T g22 = c.sqH1H2;
    T g33 = fm_div(-(c.H3 * c.Bp * c.Bp), c.sqH1H2) + fm_div(c.sqH1H2 * c.delta * c.sint2, c.H3);
    T g03 = fm_div(-(c.H3 * c.Bp), c.sqH1H2);
    if (M.rlog)
        g11 *= (c.r * c.r);
    g[0] = g00; g[1] = g03; g[2] = g11; g[3] = g22; g[4] = g33;
}

template <class T>
FM_HD_INLINE void rl_uu(const MetricParams &M, T u, T sint, T cost, T *g)
{
    RLCommon<T> c;
    rl_common(M, u, sint, cost, c);
    T g00 = fm_div(fm_div(c.H3 * c.H3 * c.Bp * c.Bp, c.sint2) - c.H1 * c.H2 * c.delta, c.sqH1H2 * c.H3 * c.delta);
    T g11 = fm_div(c.delta, c.sqH1H2);
    T g22 = fm_div(1., c.sqH1H2);
    T g33 = fm_div(c.H3, c.sqH1H2 * c.delta * c.sint2);
    T g03 = fm_div(-(c.H3 * c.Bp), c.sqH1H2 * c.delta * c.sint2);
    if (M.rlog)
        g11 *= fm_div(1.0, c.r * c.r);
    g[0] = g00; g[1] = g03; g[2] = g11; g[3] = g22; g[4] = g33;
}

// FOORT/Metric.cpp:409-440
template <class T>
FM_HD_INLINE void joh_dd(const MetricParams &M, T u, T sint, T cost, T *g)
{
    const double a = M.a;
    T r = M.rlog ? fexp(u) : u;
    T A1 = 1. + fm_div(M.a13, r * r * r);
    T A2 = 1. + fm_div(M.a22, r * r);
    T A5 = 1. + fm_div(M.a52, r * r);
    T eps_f = fm_div(M.eps3, r);
    T rho2 = r * r + a * a * cost * cost + eps_f;
    T delta = r * r + a * a - 2. * r;
    T den = fpow2((r * r + a * a) * A1 - a * a * A2 * sint * sint);
    T g00 = fm_div(-rho2 * (delta - a * a * A2 * A2 * sint * sint), den);
    T g11 = fm_div(rho2, delta * A5);
    T g22 = rho2;
    T g33 = fm_div(rho2 * sint * sint * (fpow2((r * r + a * a) * A1) - a * a * delta * sint * sint), den);
    T g03 = fm_div(-a * ((r * r + a * a) * A1 * A2 - delta) * rho2 * sint * sint, den);
    if (M.rlog)
        g11 *= (r * r);
    g[0] = g00; g[1] = g03; g[2] = g11; g[3] = g22; g[4] = g33;
}

// FOORT/Metric.cpp:442-473
template <class T>
FM_HD_INLINE void joh_uu(const MetricParams &M, T u, T sint, T cost, T *g)
{
    const double a = M.a;
    T r = M.rlog ? fexp(u) : u;
    T A1 = 1. + fm_div(M.a13, r * r * r);
    T A2 = 1. + fm_div(M.a22, r * r);
    T A5 = 1. + fm_div(M.a52, r * r);
    T eps_f = fm_div(M.eps3, r);
    T rho2 = r * r + a * a * cost * cost + eps_f;
    T delta = r * r + a * a - 2. * r;
    T g00 = fm_div(-1. * fpow2((r * r + a * a) * A1) + a * a * delta * sint * sint, delta * rho2);
    T g11 = fm_div(delta * A5, rho2);
    T g22 = fm_div(1., rho2);
    T g33 = fm_div(-a * a * A2 * A2 * sint * sint + delta, delta * rho2 * sint * sint);
    T g03 = fm_div(-a * (A2 * A1 * (r * r + a * a) - delta), delta * rho2);
    if (M.rlog)
        g11 *= fm_div(1.0, r * r);
    g[0] = g00; g[1] = g03; g[2] = g11; g[3] = g22; g[4] = g33;
}

// FOORT/Metric.cpp:516-554 / :572-610
template <class T>
struct MNCommon {
    T r, sint, f, exp_2gamma, omega, rho_sq, delta;
};

// The compiler's own divisions here and in mn_dd_from / mn_uu_from: with fm_div the shared
// reciprocals of R, R^2, ... took the finite-difference kernel to 508 of its 512 registers
// and made it 12% slower (measured), for the same bits.
template <class T>
FM_HD_INLINE void mn_common(const MetricParams &M, T u, T sint, T cost, MNCommon<T> &c)
{
    const double al = M.mn_alpha, kk = M.mn_k, a3 = M.alpha3;
    T r = M.rlog ? fexp(u) : u;
    T xx = (r - 1.) / kk;
    T yy = cost;
    T R = fsqrt(xx * xx + yy * yy - 1.);
    T P1 = xx * yy / R;
    T P2 = 0.5 * (3. * xx * xx * yy * yy / (R * R) - 1.);
    T P3 = 0.5 * (5. * fpow3(xx * yy / R) - 3. * xx * yy / R);
    T P4 = 0.125 * (35. * fpow4(xx * yy / R) - 30. * xx * xx * yy * yy / (R * R) + 3.);
    T aa = -al * fexp(2. * a3 * (1. - (xx - yy) * (1. / R + P1 / (R * R) + P2 / (R * R * R) + P3 / (R * R * R * R))));
    T bb = al * fexp(2. * a3 * (-1. + (xx + yy) * (1. / R - P1 / (R * R) + P2 / (R * R * R) - P3 / (R * R * R * R))));
    T AA = (xx * xx - 1.) * (1. + aa * bb) * (1. + aa * bb) - (1. - yy * yy) * (bb - aa) * (bb - aa);
    T BB = fpow2((xx + 1. + (xx - 1.) * aa * bb)) + fpow2(((1. + yy) * aa + (1. - yy) * bb));
    T CC = (xx * xx - 1.) * (1. + aa * bb) * (bb - aa - yy * (aa + bb)) +
                (1. - yy * yy) * (bb - aa) * (1. + aa * bb + xx * (1. - aa * bb));
    T psi = a3 * P3 / (R * R * R * R);
    T gamma_prime = 2. * a3 * a3 * (P4 * P4 - P3 * P3) / (fpow8(R)) +
                         2. * a3 * (-yy + xx * P1 / R - yy * P2 / (R * R) + xx * P3 / (R * R * R)) / R;
    c.f = fexp(2. * psi) * AA / BB;
    c.exp_2gamma = fexp(2. * gamma_prime) * AA / ((xx * xx - yy * yy) * (1. - al * al) * (1. - al * al));
    c.omega = 2. * kk * fexp(-2. * psi) * CC / AA - 4. * kk * al / (1. - al * al);
    c.rho_sq = (r - 1.) * (r - 1.) - kk * kk * cost * cost;
    c.delta = (r - 1.) * (r - 1.) - kk * kk;
    c.r = r;
    c.sint = sint;
}

template <class T>
FM_HD_INLINE void mn_dd_from(const MetricParams &M, const MNCommon<T> &c, T sint, T *g)
{
    T g00 = -c.f;
    T g22 = c.exp_2gamma * c.rho_sq / c.f;
    T g11 = g22 / c.delta;
    T g33 = -c.f * c.omega * c.omega + c.delta * sint * sint / c.f;
    T g03 = c.omega * c.f;
    if (M.rlog)
        g11 *= (c.r * c.r);
    g[0] = g00; g[1] = g03; g[2] = g11; g[3] = g22; g[4] = g33;
}

// The same potentials, restructured for the analytic formulations (not the reference's
// arithmetic): one reciprocal of R for all its powers, the Legendre polynomials of
// w = xy/R from w and w^2, the two series in 1/R in Horner form, exp(-2 psi) as
// 1 / exp(2 psi).  Algebraically identical to mn_common.
template <class T>
FM_HD_INLINE void mn_common_an(const MetricParams &M, T u, T sint, T cost, MNCommon<T> &c)
{
    const double al = M.mn_alpha, kk = M.mn_k, a3 = M.alpha3;
    T r = M.rlog ? fexp(u) : u;
    T xx = (r - 1.) * (1. / kk);
    T yy = cost;
    T x2 = xx * xx, y2 = yy * yy;
    T R = fsqrt(x2 + y2 - 1.);
    T iR = 1. / R;
    T w = xx * yy * iR;
    T w2 = w * w;
    T P1 = w;
    T P2 = 0.5 * (3. * w2 - 1.);
    T P3 = 0.5 * w * (5. * w2 - 3.);
    T P4 = 0.125 * ((35. * w2 - 30.) * w2 + 3.);
    // 1/R + P1/R^2 + P2/R^3 + P3/R^4 and 1/R - P1/R^2 + P2/R^3 - P3/R^4, i.e. (even +- odd) / R
    // with even = 1 + P2/R^2 and odd = (P1 + P3/R^2)/R
    T iR2 = iR * iR;
    // the Legendre terms die here: psi and gamma' (and their exponentials) before aa, bb
    T iR4 = iR2 * iR2;
    T psi = a3 * P3 * iR4;
    T gamma_prime = 2. * a3 * a3 * (P4 * P4 - P3 * P3) * (iR4 * iR4) +
                    2. * a3 * (-yy + (xx * P1 + (-yy * P2 + xx * P3 * iR) * iR) * iR) * iR;
    T even = P2 * iR2 + 1.;
    T odd = (P3 * iR2 + P1) * iR;
    T s_m = (even + odd) * iR;
    T s_p = (even - odd) * iR;
    T e2psi = fexp(2. * psi);
    T e2gp = fexp(2. * gamma_prime);
    T aa = -al * fexp(2. * a3 * (1. - (xx - yy) * s_m));
    T bb = al * fexp(2. * a3 * (-1. + (xx + yy) * s_p));
    T ab = aa * bb;
    T x2m1 = x2 - 1., omy2 = 1. - y2;
    T bma = bb - aa;
    T AA = x2m1 * fpow2(1. + ab) - omy2 * fpow2(bma);
    T BB = fpow2(xx + 1. + (xx - 1.) * ab) + fpow2((1. + yy) * aa + (1. - yy) * bb);
    T CC = x2m1 * (1. + ab) * (bma - yy * (aa + bb)) + omy2 * bma * (1. + ab + xx * (1. - ab));
    const double ial2 = 1. / ((1. - al * al) * (1. - al * al));
    c.f = e2psi * AA / BB;
    c.exp_2gamma = e2gp * AA * ial2 / (x2 - y2);
    c.omega = 2. * kk * CC / (e2psi * AA) - 4. * kk * al / (1. - al * al);
    c.rho_sq = (r - 1.) * (r - 1.) - kk * kk * cost * cost;
    c.delta = (r - 1.) * (r - 1.) - kk * kk;
    c.r = r;
    c.sint = sint;
}

template <class T>
FM_HD_INLINE void mn_dd(const MetricParams &M, T u, T sint, T cost, T *g)
{
    MNCommon<T> c;
    mn_common(M, u, sint, cost, c);
    mn_dd_from(M, c, sint, g);
}

template <class T>
FM_HD_INLINE void mn_uu_from(const MetricParams &M, const MNCommon<T> &c, T sint, T *g)
{
    T beta_d = c.omega * c.f;
    T gamma_dd = -c.f * c.omega * c.omega + c.delta * sint * sint / c.f;
    T beta_u = beta_d / gamma_dd;
    T alpha_sq = beta_u * beta_d + c.f;
    T g00 = -1. / alpha_sq;
    T g22 = c.f / (c.exp_2gamma * c.rho_sq);
    T g11 = c.delta * g22;
    T g33 = 1. / gamma_dd - beta_u * beta_u / alpha_sq;
    T g03 = beta_u / alpha_sq;
    if (M.rlog)
        g11 *= 1.0 / (c.r * c.r);
    g[0] = g00; g[1] = g03; g[2] = g11; g[3] = g22; g[4] = g33;
}

// mn_uu_from for the analytic formulations: the same quantities with five reciprocals
// (arcp, not correctly rounded) in place of eight divisions.
template <class T>
FM_HD_INLINE void mn_uu_from_an(const MetricParams &M, const MNCommon<T> &c, T sint, T *g)
{
    const T inv_f = arcp(c.f);
    const T beta_d = c.omega * c.f;
    const T gamma_dd = -c.f * c.omega * c.omega + c.delta * sint * sint * inv_f;
    const T inv_g = arcp(gamma_dd);
    const T beta_u = beta_d * inv_g;
    const T alpha_sq = beta_u * beta_d + c.f;
    const T inv_a = arcp(alpha_sq);
    T g00 = -inv_a;
    T g22 = c.f * arcp(c.exp_2gamma * c.rho_sq);
    T g11 = c.delta * g22;
    T g33 = inv_g - beta_u * beta_u * inv_a;
    T g03 = beta_u * inv_a;
    if (M.rlog)
        g11 *= arcp(c.r * c.r);
    g[0] = g00; g[1] = g03; g[2] = g11; g[3] = g22; g[4] = g33;
}

template <class T>
FM_HD_INLINE void mn_uu(const MetricParams &M, T u, T sint, T cost, T *g)
{
    MNCommon<T> c;
    mn_common(M, u, sint, cost, c);
    mn_uu_from(M, c, sint, g);
}

template <int METRIC, class T>
FM_HD_INLINE void metric_dd(const MetricParams &M, T u, T sint, T cost, T *g)
{
    if constexpr (METRIC == FOORT_METRIC_KERR) kerr_dd(M, u, sint, cost, g);
    else if constexpr (METRIC == FOORT_METRIC_FLATSPACE) flat_dd(M, u, sint, cost, g);
    else if constexpr (METRIC == FOORT_METRIC_RASHEEDLARSEN) rl_dd(M, u, sint, cost, g);
    else if constexpr (METRIC == FOORT_METRIC_JOHANNSEN) joh_dd(M, u, sint, cost, g);
    else mn_dd(M, u, sint, cost, g);
}

template <int METRIC, class T>
FM_HD_INLINE void metric_uu(const MetricParams &M, T u, T sint, T cost, T *g)
{
    if constexpr (METRIC == FOORT_METRIC_KERR) kerr_uu(M, u, sint, cost, g);
    else if constexpr (METRIC == FOORT_METRIC_FLATSPACE) flat_uu(M, u, sint, cost, g);
    else if constexpr (METRIC == FOORT_METRIC_RASHEEDLARSEN) rl_uu(M, u, sint, cost, g);
    else if constexpr (METRIC == FOORT_METRIC_JOHANNSEN) joh_uu(M, u, sint, cost, g);
    else mn_uu(M, u, sint, cost, g);
}

// runtime dispatch (host camera set-up only)
FM_HD_INLINE void metric_dd_rt(int id, const MetricParams &M, double u, double sint, double cost, double *g)
{
    switch (id) {
    case FOORT_METRIC_KERR: kerr_dd(M, u, sint, cost, g); break;
    case FOORT_METRIC_FLATSPACE: flat_dd(M, u, sint, cost, g); break;
    case FOORT_METRIC_RASHEEDLARSEN: rl_dd(M, u, sint, cost, g); break;
    case FOORT_METRIC_JOHANNSEN: joh_dd(M, u, sint, cost, g); break;
    default: mn_dd(M, u, sint, cost, g); break;
    }
}

FM_HD_INLINE void metric_uu_rt(int id, const MetricParams &M, double u, double sint, double cost, double *g)
{
    switch (id) {
    case FOORT_METRIC_KERR: kerr_uu(M, u, sint, cost, g); break;
    case FOORT_METRIC_FLATSPACE: flat_uu(M, u, sint, cost, g); break;
    case FOORT_METRIC_RASHEEDLARSEN: rl_uu(M, u, sint, cost, g); break;
    case FOORT_METRIC_JOHANNSEN: joh_uu(M, u, sint, cost, g); break;
    default: mn_uu(M, u, sint, cost, g); break;
    }
}

#include "contraction.inc"

// ---------------------------------------------------------------------------
// Geodesic right-hand side with finite-difference Christoffels
// (Metric::getChristoffel_udd + geoRHS): d/dlambda v = -Gamma v v
// ---------------------------------------------------------------------------
template <int METRIC>
FM_HD_INLINE void rhs_fd(const RunConst &R, const double *p, const double *v, double *out)
{
    const double u = p[1], th = p[2], h = R.h_fd;
    double s0, c0, sp, cp, sm, cm;
    // sin and cos at theta + h and theta - h shifted from theta's (one reduction, one pair of
    // polynomials: the same functions as the oracle's)
    fm_sincos3(th, th + h, th - h, &s0, &c0, &sp, &cp, &sm, &cm);
    double gp[5], gm[5], d1[5], d2[5], uu[5];
    metric_dd<METRIC, double>(R.M, u + h, s0, c0, gp);
    metric_dd<METRIC, double>(R.M, u - h, s0, c0, gm);
#pragma unroll
    for (int k = 0; k < 5; ++k)
        d1[k] = (gp[k] - gm[k]) * R.inv2h;
    metric_dd<METRIC, double>(R.M, u, sp, cp, gp);
    metric_dd<METRIC, double>(R.M, u, sm, cm, gm);
#pragma unroll
    for (int k = 0; k < 5; ++k)
        d2[k] = (gp[k] - gm[k]) * R.inv2h;
    metric_uu<METRIC, double>(R.M, u, s0, c0, uu);
    geo_contract<double>(d1, d2, uu, v, out);
}

// Geodesic right-hand side from the first derivatives of g_dd and from g^uu,
// in Euler-Lagrange form: g_mb a^b = 1/2 d_m g_ab v^a v^b - d_a g_mb v^a v^b,
// with derivatives along the radial coordinate (d1) and theta (d2) only.
// Used by the analytic formulations (not the reference's summation order).
template <class S>
FM_HD_INLINE void kj_contract(const S *d1, const S *d2, const S *u, const S *v, S *out)
{
    // D*_m = d*g_mb v^b
    const S r_t = d1[0] * v[0] + d1[1] * v[3];
    const S r_r = d1[2] * v[1];
    const S r_h = d1[3] * v[2];
    const S r_p = d1[1] * v[0] + d1[4] * v[3];
    const S h_t = d2[0] * v[0] + d2[1] * v[3];
    const S h_r = d2[2] * v[1];
    const S h_h = d2[3] * v[2];
    const S h_p = d2[1] * v[0] + d2[4] * v[3];
    const S Kr = S(0.5) * (r_t * v[0] + r_r * v[1] + r_h * v[2] + r_p * v[3]);
    const S Kh = S(0.5) * (h_t * v[0] + h_r * v[1] + h_h * v[2] + h_p * v[3]);
    const S Ft = -(r_t * v[1] + h_t * v[2]);
    const S Fr = Kr - (r_r * v[1] + h_r * v[2]);
    const S Fh = Kh - (r_h * v[1] + h_h * v[2]);
    const S Fp = -(r_p * v[1] + h_p * v[2]);
    out[0] = u[0] * Ft + u[1] * Fp;
    out[1] = u[2] * Fr;
    out[2] = u[3] * Fh;
    out[3] = u[1] * Ft + u[4] * Fp;
}

FM_HD_INLINE float srcp(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}
FM_HD_INLINE double srcp(double x) { return fm_div(1.0, x); }

// Kerr (M = 1) in closed form: the first derivatives of the Boyer-Lindquist g_dd of
// FOORT/Metric.cpp:131-159 and g^uu of :162-191, written with Sigma = r^2 + a^2 cos^2,
// Delta = r^2 - 2r + a^2, q = Sigma - 2 r^2, in the Euler-Lagrange form of kj_contract:
//   g_mb a^b = F_m = 1/2 d_m g(v, v) - d_a g_mb v^a v^b   (a, m over r and theta).
// With e = half the derivatives (e1 along r, e2 along theta) and A_m = e1_mb v^b,
// B_m = e2_mb v^b:
//   -F_t / 2 = A_t v^r + B_t v^th,   -F_ph / 2 = A_ph v^r + B_ph v^th,
//   F_r  = A_t v^t + A_ph v^ph + (e1_thth v^th - 2 e2_rr v^r) v^th - e1_rr (v^r)^2,
//   F_th = B_t v^t + B_ph v^ph + (e2_rr v^r - 2 e1_thth v^th) v^r - e2_thth (v^th)^2,
// where the t and phi rows factor:
//   A_t = q / Sigma^2 (v^t - a s^2 v^ph),   A_ph = s^2 (r v^ph - a A_t),
//   B_t = k (a v^t - (r^2 + a^2) v^ph),     k = 2 r a s c / Sigma^2,
//   B_ph = s c (r^2 + a^2) v^ph + k (a s^2 (2 Sigma + a^2 s^2) v^ph - (r^2 + a^2) v^t),
// (s = sin theta, c = cos theta; e1_thth = r, e2_thth = -a^2 s c, e2_rr = -a^2 s c / Delta,
// e1_rr = (r Delta - Sigma (r - 1)) / Delta^2), and a = g^uu F with the factor -2 of the t
// and phi rows folded into w0 = -2 g^tt, w1 = -2 g^tph, w4 = -2 g^phph.  With RLogScale the
// radial coordinate is u = log r: v^r = r v^u and a^u = a^r / r - (v^u)^2.
// RL: 0 or 1 fixes RLogScale at compile time (the step then carries no select for the
// other case), -1 reads it from R.
template <class S, class T, int RL = -1>
FM_HD_INLINE void rhs_kerr_closed_sc(const RunConst &R, const T *p, const T *v, S s, S c, T *out)
{
    const bool rlog = RL < 0 ? R.M.rlog != 0 : RL != 0;
    const S a = static_cast<S>(R.M.a);
    const S a2 = a * a;
    S r, vr;
    if (rlog) {
        r = sexp(static_cast<S>(p[1]));
        vr = r * static_cast<S>(v[1]);
    } else {
        r = static_cast<S>(p[1]);
        vr = static_cast<S>(v[1]);
    }
    const S vt = static_cast<S>(v[0]), vh = static_cast<S>(v[2]), vp = static_cast<S>(v[3]);
    const S r2 = r * r, s2 = s * s, sc = s * c;
    const S rpa = r2 + a2;
    const S Sig = r2 + a2 * (c * c);
    const S Del = rpa - S(2) * r;
    // one reciprocal for 1/Sigma, 1/Delta and 1/sin^2
    const S SD = Sig * Del;
    const S inv = srcp(SD * s2);
    const S iSD = s2 * inv;
    const S iS = Del * iSD, iD = Sig * iSD, is2 = SD * inv;
    const S iS2 = iS * iS;
    const S qS = (Sig - S(2) * r2) * iS2;
    const S as2 = a * s2, a2s2 = a2 * s2, a2sc = a2 * sc;
    // the t and phi rows
    const S At = qS * (vt - as2 * vp);
    const S Ap = s2 * (r * vp - a * At);
    const S k = ((r + r) * iS2) * (a * sc);
    const S Bt = k * (a * vt - rpa * vp);
    const S Bp = (sc * rpa) * vp + k * ((as2 * (S(2) * Sig + a2s2)) * vp - rpa * vt);
    // the r and theta rows
    const S e1rr = (r * Del - Sig * (r - S(1))) * (iD * iD);
    const S P = r * vh, Q = -(a2sc * iD) * vr;
    const S Ft = At * vr + Bt * vh;
    const S Fp = Ap * vr + Bp * vh;
    const S Fr = At * vt + Ap * vp + (P - S(2) * Q) * vh - e1rr * (vr * vr);
    const S Fh = Bt * vt + Bp * vp + (Q - S(2) * P) * vr + a2sc * (vh * vh);
    const S iSD2 = iSD + iSD;
    const S w0 = (rpa * rpa - a2s2 * Del) * iSD2;
    const S w1 = ((a + a) * r) * iSD2;
    const S w4 = (a2s2 - Del) * (iSD2 * is2);
    out[0] = static_cast<T>(w0 * Ft + w1 * Fp);
    out[2] = static_cast<T>(iS * Fh);
    out[3] = static_cast<T>(w1 * Ft + w4 * Fp);
    const S o1 = (Del * iS) * Fr;
    if (rlog) {
        const T vu = v[1];
        out[1] = static_cast<T>(o1 * srcp(r)) - vu * vu;
    } else {
        out[1] = static_cast<T>(o1);
    }
}

template <class S, class T, int RL = -1>
FM_HD_INLINE void rhs_kerr_closed(const RunConst &R, const T *p, const T *v, T *out)
{
    S s, c;
    ssincos(p[2], &s, &c);
    rhs_kerr_closed_sc<S, T, RL>(R, p, v, s, c, out);
}

// Analytic right-hand side: the metric and its exact derivatives in one
// dual-number pass (scalar type S), then the same sparse contraction.  s, c: sin and cos
// of p[2].
template <int METRIC, class S, class T>
FM_HD_INLINE void rhs_an_sc(const RunConst &R, const T *p, const T *v, S s, S c, T *out)
{
    const D2<S> U(static_cast<S>(p[1]), S(1), S(0));
    const D2<S> ST(s, S(0), c), CT(c, S(0), -s);
    D2<S> g[5];
    S d1[5], d2[5], uu[5], vv[4], o[4];
    if constexpr (METRIC == FOORT_METRIC_MANKONOVIKOV) {
        // g_dd and g^uu share one evaluation of the Manko-Novikov potentials
        // (their most expensive part); g^uu needs only its values
        MNCommon<D2<S>> cd;
        mn_common_an(R.M, U, ST, CT, cd);
        mn_dd_from(R.M, cd, ST, g);
        const MNCommon<S> cv{cd.r.v, cd.sint.v, cd.f.v, cd.exp_2gamma.v, cd.omega.v, cd.rho_sq.v, cd.delta.v};
        mn_uu_from_an(R.M, cv, s, uu);
    } else {
        D2<S> gu[5];
        metric_dd<METRIC, D2<S>>(R.M, U, ST, CT, g);
        metric_uu<METRIC, D2<S>>(R.M, U, ST, CT, gu);
#pragma unroll
        for (int k = 0; k < 5; ++k)
            uu[k] = gu[k].v;
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        d1[k] = g[k].du;
        d2[k] = g[k].dt;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
        vv[k] = static_cast<S>(v[k]);
    kj_contract<S>(d1, d2, uu, vv, o);
#pragma unroll
    for (int k = 0; k < 4; ++k)
        out[k] = static_cast<T>(o[k]);
}

template <int METRIC, class S, class T>
FM_HD_INLINE void rhs_an(const RunConst &R, const T *p, const T *v, T *out)
{
    S s, c;
    ssincos(p[2], &s, &c);
    rhs_an_sc<METRIC, S, T>(R, p, v, s, c, out);
}

// Right-hand-side policies: FD = the reference's finite differences (exact
// parity), AN64 = analytic binary64, MIXED = analytic, binary32 away from the
// strong-field region (u > mixed_u) and the poles, binary64 elsewhere.
template <int METRIC>
struct RhsFD {
    static constexpr int metric = METRIC;
    static constexpr bool exact = true;
    static constexpr bool split_rk4 = false;
    static constexpr bool stage_sc = false;
    FM_HD_INLINE static void eval(const RunConst &R, const double *p, const double *v, double *out)
    {
        rhs_fd<METRIC>(R, p, v, out);
    }
};

template <int METRIC, class S, class T = double, int RL = -1>
FM_HD_INLINE void rhs_analytic(const RunConst &R, const T *p, const T *v, T *out)
{
    if constexpr (METRIC == FOORT_METRIC_KERR)
        rhs_kerr_closed<S, T, RL>(R, p, v, out);
    else
        rhs_an<METRIC, S, T>(R, p, v, out);
}

template <int METRIC, class S, class T = S, int RL = -1>
FM_HD_INLINE void rhs_analytic_sc(const RunConst &R, const T *p, const T *v, S s, S c, T *out)
{
    if constexpr (METRIC == FOORT_METRIC_KERR)
        rhs_kerr_closed_sc<S, T, RL>(R, p, v, s, c, out);
    else
        rhs_an_sc<METRIC, S, T>(R, p, v, s, c, out);
}

// sin and cos of th0 + d from s0 = sin th0, c0 = cos th0, for |d| <= 0.1: Taylor
// polynomials of sin d (to d^9) and cos d - 1 (to d^10), truncation below 3e-18 relative,
// joined by the addition formulas.  Within a few ulp of sin and cos of th0 + d.
FM_HD_INLINE void sincos_shift(double s0, double c0, double d, double *s, double *c)
{
    const double z = d * d;
    const double sd = d + (d * z) * (-1.0 / 6.0 + z * (1.0 / 120.0 + z * (-1.0 / 5040.0 + z * (1.0 / 362880.0))));
    const double cm1 = z * (-0.5 + z * (1.0 / 24.0 + z * (-1.0 / 720.0 + z * (1.0 / 40320.0 + z * (-1.0 / 3628800.0)))));
    *s = s0 + (s0 * cm1 + c0 * sd);
    *c = c0 + (c0 * cm1 - s0 * sd);
}

// sin and cos of a stage's theta th from those of the step's starting theta th0: the
// shift where the stage is close (the adaptive step keeps |th - th0| near eps min(th0,
// pi - th0)), the full evaluation otherwise
FM_HD_INLINE void stage_sincos(double s0, double c0, double th0, double th, double *s, double *c)
{
    const double d = th - th0;
    if (__builtin_fabs(d) <= 0.1)
        sincos_shift(s0, c0, d, s, c);
    else
        ssincos(th, s, c);
}

// The analytic policies take RL as the right-hand sides do (-1: RLogScale read from the
// run constants); Fixed<RL> is the same policy with RLogScale fixed, which geo_update
// selects once per step for the closed-form Kerr right-hand side.
template <int METRIC, int RL = -1>
struct RhsAn64 {
    static constexpr int metric = METRIC;
    static constexpr bool exact = false;
    static constexpr bool split_rk4 = false;
    static constexpr bool stage_sc = true; // RK4 stages take sin/cos of theta from the step's start
    static constexpr bool fix_rlog = METRIC == FOORT_METRIC_KERR && RL < 0;
    template <int X>
    using Fixed = RhsAn64<METRIC, X>;
    FM_HD_INLINE static void eval(const RunConst &R, const double *p, const double *v, double *out)
    {
        rhs_analytic<METRIC, double, double, RL>(R, p, v, out);
    }
    FM_HD_INLINE static void eval_sc(const RunConst &R, const double *p, const double *v, double s, double c,
                                     double *out)
    {
        rhs_analytic_sc<METRIC, double, double, RL>(R, p, v, s, c, out);
    }
};

template <int METRIC, int RL = -1>
struct RhsMixed {
    static constexpr int metric = METRIC;
    static constexpr bool exact = false;
    static constexpr bool split_rk4 = true; // RK4 steps go through rk4_step_mixed
    static constexpr bool stage_sc = false;
    static constexpr bool fix_rlog = METRIC == FOORT_METRIC_KERR && RL < 0;
    template <int X>
    using Fixed = RhsMixed<METRIC, X>;
    using Near = RhsAn64<METRIC, RL>;       // the binary64 step near the hole and the poles
    FM_HD_INLINE static bool far(const RunConst &R, const double *p)
    {
        const double th = p[2];
        return p[1] > R.mixed_u && th > R.mixed_pole && (kPi - th) > R.mixed_pole;
    }
    // the binary32 right-hand side of the far field
    FM_HD_INLINE static void eval_far(const RunConst &R, const float *p, const float *v, float *out)
    {
        rhs_analytic<METRIC, float, float, RL>(R, p, v, out);
    }
    FM_HD_INLINE static void eval(const RunConst &R, const double *p, const double *v, double *out)
    {
        if (far(R, p))
            rhs_analytic<METRIC, float, double, RL>(R, p, v, out);
        else
            rhs_analytic<METRIC, double, double, RL>(R, p, v, out);
    }
};

// std::min / std::max semantics
FM_HD_INLINE double ref_min(double a, double b) { return (b < a) ? b : a; }
FM_HD_INLINE double ref_max(double a, double b) { return (a < b) ? b : a; }

// Integrators::GetAdaptiveStep, FOORT/Integrators.cpp:34-47
FM_HD_INLINE double adaptive_step(const RunConst &R, const double *pos, const double *vel)
{
    double d1 = fm_div(R.eps, __builtin_fabs(vel[1]) + 1e-20);
    double d2 = fm_div(R.eps * ref_min(pos[2], kPi - pos[2]), __builtin_fabs(vel[2]) + 1e-20);
    double d3 = fm_div(R.eps, __builtin_fabs(vel[3] + 1e-20));
    double h = fm_div(1.0, fm_div(1.0, __builtin_fabs(d1)) + fm_div(1.0, __builtin_fabs(d2)) +
                               fm_div(1.0, __builtin_fabs(d3)));
    return ref_max(h, R.minstep);
}

// The same step size in an equivalent form with two divisions instead of six,
// h = eps / ((|v1| + d) + (|v2| + d) / min(th, pi - th) + |v3 + d|), for the
// formulations that are not bit-compatible with the reference anyway.
FM_HD_INLINE double adaptive_step_fast(const RunConst &R, const double *pos, const double *vel)
{
    // eps / (|v1|+d + (|v2|+d)/m + |v3+d|), multiplied through by m: one division (at the
    // poles, m = 0, both forms give 0 and so the smallest step)
    const double m = ref_min(pos[2], kPi - pos[2]);
    const double den = ((__builtin_fabs(vel[1]) + 1e-20) + __builtin_fabs(vel[3] + 1e-20)) * m +
                       (__builtin_fabs(vel[2]) + 1e-20);
    // the compiler's own division here: this quotient heads every step's chain of dependent
    // instructions, and without v_div_scale_f64 between v_rcp_f64 and its first use the
    // mixed kernel ran 5% slower on C1 (measured), for the same bits
    return ref_max(R.eps * m / den, R.minstep);
}

template <class RHS>
FM_HD_INLINE double step_size(const RunConst &R, const double *pos, const double *vel)
{
    if constexpr (RHS::exact)
        return adaptive_step(R, pos, vel);
    else
        return adaptive_step_fast(R, pos, vel);
}

// IntegrateGeodesicStep_RK4, FOORT/Integrators.cpp:51-94
template <class RHS>
FM_HD_INLINE void rk4_step(const RunConst &R, const double *x, const double *v, double *nx, double *nv, double &hout)
{
    // The sums ((k1 + 2 k2) + 2 k3) + k4 and ((v + 2 l2) + 2 l3) + l4 are accumulated as the
    // stages come (the same additions in the same order, so the same bits): k1..k3 and l2, l3
    // need not stay live across the later right-hand sides.
    const double h = step_size<RHS>(R, x, v);
    const double hh = 0.5 * h;
    double k[4], sv[4], sx[4], l[4], tp[4];
    // the analytic policies take sin/cos of theta: evaluated once at the step's start and
    // shifted to the later stages (stage_sincos)
    double s0 = 0, c0 = 0;
    auto eval = [&](const double *pp, const double *vv, double *out) {
        if constexpr (RHS::stage_sc) {
            double s, c;
            stage_sincos(s0, c0, x[2], pp[2], &s, &c);
            RHS::eval_sc(R, pp, vv, s, c, out);
        } else {
            RHS::eval(R, pp, vv, out);
        }
    };
    if constexpr (RHS::stage_sc) {
        ssincos(x[2], &s0, &c0);
        RHS::eval_sc(R, x, v, s0, c0, k);
    } else {
        RHS::eval(R, x, v, k);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        tp[i] = x[i] + v[i] * hh;
        l[i] = v[i] + k[i] * hh; // l2
        sv[i] = k[i];
        sx[i] = v[i] + l[i] * 2;
    }
    eval(tp, l, k);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        tp[i] = x[i] + l[i] * hh;
        l[i] = v[i] + k[i] * hh; // l3
        sv[i] = sv[i] + k[i] * 2;
        sx[i] = sx[i] + l[i] * 2;
    }
    eval(tp, l, k);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        tp[i] = x[i] + l[i] * h;
        l[i] = v[i] + k[i] * h; // l4
        sv[i] = sv[i] + k[i] * 2;
        sx[i] = sx[i] + l[i];
    }
    eval(tp, l, k);
    // the reference divides (FOORT/Integrators.cpp:91-92); the analytic formulations multiply by
    // the reciprocal, one instruction instead of a correctly rounded division
    const double h6 = RHS::exact ? fm_div(h, 6.0) : h * (1.0 / 6.0);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        nv[i] = v[i] + (sv[i] + k[i]) * h6;
        nx[i] = x[i] + sx[i] * h6;
    }
    hout = h;
}

// RK4 step of the mixed policy.  The precision is chosen once per step, from
// the step's starting point.  In the far field all four stages run in binary32
// and only their increments reach the binary64 state, combined as
//   v' = v + (h/6)(k1 + 2 k2 + 2 k3 + k4),   x' = x + h v + (h^2/6)(k1 + k2 + k3),
// which is the RK4 combination above (l2 - v = (h/2) k1, l3 - v = (h/2) k2,
// l4 - v = h k3).  The step size stays binary64.  Near the hole and the poles
// the step is the binary64 analytic one.
template <class RHS>
FM_HD_INLINE void rk4_step_mixed(const RunConst &R, const double *x, const double *v, double *nx, double *nv,
                                 double &hout)
{
    if (!RHS::far(R, x)) {
        rk4_step<typename RHS::Near>(R, x, v, nx, nv, hout);
        return;
    }
    const double h = adaptive_step_fast(R, x, v);
    const float hf = static_cast<float>(h), hh = 0.5f * hf;
    // sums accumulated as the stages come, in the order of the formula (as in rk4_step)
    float xf[4], vf[4], k[4], l[4], tp[4], sv[4], sk[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        xf[i] = static_cast<float>(x[i]);
        vf[i] = static_cast<float>(v[i]);
    }
    RHS::eval_far(R, xf, vf, k);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        tp[i] = xf[i] + vf[i] * hh;
        l[i] = vf[i] + k[i] * hh; // l2
        sv[i] = k[i];
        sk[i] = k[i];
    }
    RHS::eval_far(R, tp, l, k);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        tp[i] = xf[i] + l[i] * hh;
        l[i] = vf[i] + k[i] * hh; // l3
        sv[i] = sv[i] + k[i] * 2.0f;
        sk[i] = sk[i] + k[i];
    }
    RHS::eval_far(R, tp, l, k);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        tp[i] = xf[i] + l[i] * hf;
        l[i] = vf[i] + k[i] * hf; // l4
        sv[i] = sv[i] + k[i] * 2.0f;
        sk[i] = sk[i] + k[i];
    }
    RHS::eval_far(R, tp, l, k);
    const float h6 = static_cast<float>(h * (1.0 / 6.0)), hsq6 = static_cast<float>(h * h * (1.0 / 6.0));
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        nv[i] = v[i] + static_cast<double>((sv[i] + k[i]) * h6);
        nx[i] = (x[i] + v[i] * h) + static_cast<double>(sk[i] * hsq6);
    }
    hout = h;
}

// Verlet iterations are capped (the reference loops without bound); the CPU
// oracle uses the same cap.
constexpr int kVerletMaxIter = 1000;

// IntegrateGeodesicStep_Verlet, FOORT/Integrators.cpp:99-152
template <class RHS>
FM_HD_INLINE void verlet_step(const RunConst &R, const double *x, const double *v, double *nx, double *nv, double &hout)
{
    const double h = step_size<RHS>(R, x, v);
    double ac[4], as[4], vi[4];
    // the analytic policies take sin/cos of theta: evaluated at x, shifted to nx (stage_sincos),
    // and kept for the velocity iterations at nx
    double s1 = 0, c1 = 0;
    auto eval_nx = [&](const double *vv, double *out) {
        if constexpr (RHS::stage_sc)
            RHS::eval_sc(R, nx, vv, s1, c1, out);
        else
            RHS::eval(R, nx, vv, out);
    };
    if constexpr (RHS::stage_sc) {
        double s0, c0;
        ssincos(x[2], &s0, &c0);
        RHS::eval_sc(R, x, v, s0, c0, ac);
        const double hh2 = h * h / 2.0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            nx[i] = (x[i] + v[i] * h) + ac[i] * hh2;
            vi[i] = v[i] + ac[i] * h;
        }
        stage_sincos(s0, c0, x[2], nx[2], &s1, &c1);
    } else {
        RHS::eval(R, x, v, ac);
        const double hh2 = h * h / 2.0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            nx[i] = (x[i] + v[i] * h) + ac[i] * hh2;
            vi[i] = v[i] + ac[i] * h;
        }
    }
    eval_nx(vi, as);
    const double h2 = h / 2.0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        nv[i] = v[i] + (ac[i] + as[i]) * h2;
    if (R.verlet_tol > 0.0) {
        const double tol2 = R.verlet_tol * R.verlet_tol;
        for (int it = 0; it < kVerletMaxIter; ++it) {
            double n1 = 0, n2 = 0;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                double dd = nv[i] - vi[i];
                n1 += dd * dd;
                n2 += nv[i] * nv[i];
            }
            if (!(n1 / n2 > tol2))
                break;
#pragma unroll
            for (int i = 0; i < 4; ++i)
                vi[i] = nv[i];
            eval_nx(vi, as);
#pragma unroll
            for (int i = 0; i < 4; ++i)
                nv[i] = v[i] + (ac[i] + as[i]) * h2;
        }
    }
    hout = h;
}

// ---------------------------------------------------------------------------
// Geodesic state, terminations and diagnostics
// ---------------------------------------------------------------------------
struct GeoState {
    double pos[4], vel[4], lambda;
    double prevtheta, closest;
    int32_t term;
    uint32_t steps;
    uint32_t hc, bc, tc, thc, tsteps;   // termination counters
    uint32_t epc, crc;                  // diagnostic counters
    int32_t quadrant, passes;
};

// Termination::DecideUpdate, FOORT/Terminations.cpp:71-90
// EVERY: every termination and diagnostic updates on every step (UpdateFrequency 1,
// or 0 for a termination), known at compile time, so no counters are kept.
template <bool EVERY>
FM_HD_INLINE bool term_decide(uint32_t n, uint32_t &cnt)
{
    if (EVERY || n == 0)
        return true;
    ++cnt;
    if (cnt >= n) {
        cnt = 0;
        return true;
    }
    return false;
}

// Diagnostic::DecideUpdate, FOORT/Diagnostics.cpp:113-140
template <bool EVERY>
FM_HD_INLINE bool diag_decide(uint32_t n, int32_t start, int32_t finish, uint32_t &cnt, const GeoState &g)
{
    if (EVERY)
        return true;
    if (n > 0) {
        ++cnt;
        if (cnt >= n) {
            cnt = 0;
            return true;
        }
        return false;
    } else if (start && g.lambda == 0.0) {
        return true;
    } else if (finish && g.term != FOORT_TERM_CONTINUE) {
        return true;
    }
    return false;
}

// FourColorScreenDiagnostic::UpdateData, FOORT/Diagnostics.cpp:159-200
FM_HD_INLINE int32_t four_color_quadrant(const double *pos)
{
    double p3 = pos[3];
    if (__builtin_fabs(p3) < 2 * kPi * 1e5) {
        while (p3 > 2 * kPi)
            p3 -= 2 * kPi;
        while (p3 < 0)
            p3 += 2 * kPi;
    }
    if (pos[2] < kPi / 2)
        return (p3 < kPi) ? 1 : 2;
    return (p3 < kPi) ? 3 : 4;
}

// Termination and diagnostic flags read afresh where they are tested (a uniform value in a
// scalar register, tested by scalar instructions).  Tested once before the step loop, each flag
// becomes a 64-bit lane mask in a scalar register pair held across the loop; in the one-shot
// kernels those pairs spilled to vector-register lanes and were reloaded by vector
// instructions on every step.  Used by the one-shot kernels only (LIVE): measured on MI355X on
// the 1024x1024 Kerr image, the mixed one-shot kernel 4.5% faster with it, the mixed
// persistent kernel 10% slower.
FM_HD_INLINE uint32_t live_flags(uint32_t f)
{
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+s"(f));
#endif
    return f;
}

// LIVE: the flags through live_flags (the one-shot kernels' step loop, whose run constants are
// the kernel's uniform argument; never where R may be a lane's own copy, as a queue request's).
template <bool EVERY, bool LIVE = false>
FM_HD_INLINE void diag_update(const RunConst &R, GeoState &g)
{
    const uint32_t diags = LIVE ? live_flags(R.diags) : R.diags;
    if (diags & FOORT_DIAGFLAG_FOURCOLORSCREEN) {
        if (g.term == FOORT_TERM_BOUNDARYSPHERE)
            g.quadrant = four_color_quadrant(g.pos);
    }
    if (diags & FOORT_DIAGFLAG_EQUATORIALPASSES) { // FOORT/Diagnostics.cpp:351-380
        if (diag_decide<EVERY>(R.ep_n, R.ep_start, R.ep_finish, g.epc, g)) {
            double cur = g.pos[2];
            if (__builtin_fabs(cur - kPi / 2.0) > kPi / 2.0 * R.ep_thr) {
                if (g.prevtheta > 0 && ((g.prevtheta - kPi / 2.0) * (cur - kPi / 2.0)) < 0.0)
                    ++g.passes;
                g.prevtheta = cur;
            }
        }
        if (g.term == FOORT_TERM_HORIZON)
            g.passes = -g.passes;
    }
    if (diags & FOORT_DIAGFLAG_CLOSESTRADIUS) { // FOORT/Diagnostics.cpp:427-446
        if (diag_decide<EVERY>(R.cr_n, R.cr_start, R.cr_finish, g.crc, g)) {
            double cur = true_r(R.M, g.pos[1]);
            if (g.closest < 0 || cur < g.closest)
                g.closest = cur;
        }
        if (g.term == FOORT_TERM_HORIZON)
            g.closest = 0.0;
    }
}

// Terminations in the order of CreateTerminationVector, FOORT/Terminations.cpp:22-40
template <bool EVERY, bool LIVE = false>
FM_HD_INLINE int32_t check_terms(const RunConst &R, GeoState &g)
{
    const uint32_t terms = LIVE ? live_flags(R.terms) : R.terms;
    if (terms & FOORT_TERMFLAG_HORIZON) { // :96-117
        if (term_decide<EVERY>(R.h_n, g.hc)) {
            double r = true_r(R.M, g.pos[1]);
            if ((r < R.M.horizon * (1 + R.h_eps)) || (R.M.horizon == 0.0 && r < R.h_eps))
                return FOORT_TERM_HORIZON;
        }
    }
    if (terms & FOORT_TERMFLAG_BOUNDARYSPHERE) { // :130-146
        if (term_decide<EVERY>(R.s_n, g.bc)) {
            double r = true_r(R.M, g.pos[1]);
            if (r > R.sphere_R)
                return FOORT_TERM_BOUNDARYSPHERE;
        }
    }
    if (terms & FOORT_TERMFLAG_TIMEOUT) { // :168-186
        if (term_decide<EVERY>(R.t_n, g.tc)) {
            if (g.tsteps >= R.max_steps)
                return FOORT_TERM_TIMEOUT;
            ++g.tsteps;
        }
    }
    if (terms & FOORT_TERMFLAG_THETASINGULARITY) { // :200-214
        if (term_decide<EVERY>(R.th_n, g.thc)) {
            double th = g.pos[2];
            if (__builtin_fabs(th) < R.theta_eps || __builtin_fabs(kPi - th) < R.theta_eps)
                return FOORT_TERM_THETASINGULARITY;
        }
    }
    return FOORT_TERM_CONTINUE;
}

// ViewScreen::SetNewInitialConditions + Geodesic::Reset for the unit screen point (ux, uy)
template <bool EVERY>
FM_HD_INLINE void geo_init(const RunConst &R, double ux, double uy, GeoState &g)
{
    g.pos[0] = R.cam_t;
    g.pos[1] = R.cam_u;
    g.pos[2] = R.cam_theta;
    g.pos[3] = R.cam_phi;
    const double alpha = R.center_x + R.size_x * (ux - 0.5);
    const double beta = R.center_y + R.size_y * (uy - 0.5);
    const double rc = R.cam_r;
    const double densqrt = __builtin_sqrt(rc * rc + alpha * alpha + beta * beta);
    const double pf[4] = {-1.0, -rc / densqrt, -beta / rc / densqrt, alpha / rc / R.sin_theta0 / densqrt};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        double s = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            s += R.V[j][i] * pf[j];
        g.vel[i] = s;
    }
    double cur = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        cur += R.g0[i] * g.vel[i];
    const double inv = 1 / cur;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        g.vel[i] = (g.vel[i] * 1.0) * inv;
    g.lambda = 0.0;
    g.term = FOORT_TERM_CONTINUE;
    g.steps = 0;
    g.hc = g.bc = g.tc = g.thc = g.tsteps = 0;
    g.epc = g.crc = 0;
    g.quadrant = 0;
    g.passes = 0;
    g.prevtheta = -1;
    g.closest = -1;
    diag_update<EVERY>(R, g);
}

// A geodesic whose position and velocity are all NaN (one that went through a pole with
// ThetaSingularity off, say) stays NaN: every later step computes NaN from NaN, and no
// termination or diagnostic condition holds for NaN (each is a comparison, false for NaN), so
// the steps change nothing but the step counters until TimeOut counts out or the step cap is
// reached (the reference takes those steps: such a geodesic ends by TimeOut).  When every
// termination and diagnostic updates on every step, the counters' values are known in closed
// form: jump to the step before the one that ends the geodesic, which then runs as usual.
// Returns the steps skipped; the result is the one of stepping through.  (ClosestRadius: a
// NaN radius is taken only while the kept value is negative, and then once is as many times.)
FM_HD_INLINE uint32_t skip_nan_steps(const RunConst &R, GeoState &g)
{
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (!(g.pos[i] != g.pos[i]) || !(g.vel[i] != g.vel[i]))
            return 0;
    uint64_t last = g.steps + 1ull > R.step_cap ? g.steps + 1ull : R.step_cap; // STEPCAP: steps >= step_cap
    const bool timeout = (R.terms & FOORT_TERMFLAG_TIMEOUT) != 0;
    if (timeout) {
        // TimeOut ends the step on which tsteps >= max_steps holds on entry; each step before
        // it adds one to tsteps
        const uint64_t to = g.steps + 1ull + (g.tsteps >= R.max_steps ? 0ull : uint64_t(R.max_steps - g.tsteps));
        last = to < last ? to : last;
    }
    const uint32_t skip = static_cast<uint32_t>(last - 1ull - g.steps);
    g.steps += skip;
    if (timeout)
        g.tsteps += skip;
    return skip;
}

// true for a policy whose fix_rlog is true (policies without the member: false)
template <class RHS>
constexpr auto rhs_fixes_rlog_(int) -> decltype(RHS::fix_rlog, bool())
{
    return RHS::fix_rlog;
}
template <class RHS>
constexpr bool rhs_fixes_rlog_(long)
{
    return false;
}
template <class RHS>
constexpr bool rhs_fixes_rlog()
{
    return rhs_fixes_rlog_<RHS>(0);
}

template <class RHS>
FM_HD_INLINE void geo_rk4(const RunConst &R, const double *x, const double *v, double *nx, double *nv, double &h)
{
    if constexpr (RHS::split_rk4)
        rk4_step_mixed<RHS>(R, x, v, nx, nv, h);
    else
        rk4_step<RHS>(R, x, v, nx, nv, h);
}

// Geodesic::Update, FOORT/Geodesic.cpp:61-88
template <class RHS, int INTEG, bool EVERY, bool LIVE = false>
FM_HD_INLINE void geo_update(const RunConst &R, GeoState &g)
{
    double np[4], nv[4], h;
    if constexpr (INTEG == FOORT_INTEGRATOR_VERLET)
        verlet_step<RHS>(R, g.pos, g.vel, np, nv, h);
    else if constexpr (rhs_fixes_rlog<RHS>()) {
        // one uniform branch per step instead of selects in every right-hand side
        if (R.M.rlog)
            geo_rk4<typename RHS::template Fixed<1>>(R, g.pos, g.vel, np, nv, h);
        else
            geo_rk4<typename RHS::template Fixed<0>>(R, g.pos, g.vel, np, nv, h);
    } else
        geo_rk4<RHS>(R, g.pos, g.vel, np, nv, h);
    g.lambda += h;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        g.pos[i] = np[i];
        g.vel[i] = nv[i];
    }
    ++g.steps;
    int32_t t = check_terms<EVERY, LIVE>(R, g);
    if (t == FOORT_TERM_CONTINUE && g.steps >= R.step_cap)
        t = FOORT_TERM_STEPCAP;
    g.term = t;
    diag_update<EVERY, LIVE>(R, g);
}

// Called by the kernels between runs of steps (a test on every step costs the step loop a
// few percent): a geodesic whose state has turned NaN jumps to its last step.  lambda is
// NaN from the first NaN step size on, so one comparison covers the test.
template <bool EVERY>
FM_HD_INLINE void geo_skip_nan(const RunConst &R, GeoState &g)
{
    if constexpr (EVERY)
        if (g.term == FOORT_TERM_CONTINUE && g.lambda != g.lambda)
            skip_nan_steps(R, g);
}

// steps between two tests for NaN in the kernels that run a geodesic to its end
constexpr uint32_t kNanTestSteps = 64;

} // namespace foort
