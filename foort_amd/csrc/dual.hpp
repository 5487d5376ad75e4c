// dual.hpp -- forward-mode dual numbers carrying the two partial derivatives
// d/du (radial coordinate) and d/dtheta, used to evaluate a metric and its
// exact first derivatives in one pass from the same formula the
// finite-difference path uses.  S is float or double; double operands mixed
// in (the metric parameters and literal constants) are converted to S.
#pragma once

#include "fm_div.hpp"

#if defined(__HIPCC__)
#define DUAL_HD __host__ __device__ __attribute__((always_inline)) inline
#else
#define DUAL_HD inline
#endif

namespace foort {

// 1 / x for the analytic formulations, which are not bound to the reference's rounding: on the
// GPU the hardware reciprocal refined by Newton steps (two in binary64, one in binary32) to
// within about an ulp, with v_div_fixup for zeros, infinities and NaN (so 1 / 0 is still an
// infinity), instead of a correctly rounded division (9 binary64 or 10 binary32 instructions).
DUAL_HD double arcp(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    return __builtin_amdgcn_div_fixup(r, x, 1.0);
#else
    return 1.0 / x;
#endif
}
DUAL_HD float arcp(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    float r = __builtin_amdgcn_rcpf(x);
    const float e = __builtin_fmaf(-x, r, 1.0f);
    r = __builtin_fmaf(r, e, r);
    return __builtin_amdgcn_div_fixupf(r, x, 1.0f);
#else
    return 1.0f / x;
#endif
}

template <class S>
struct D2 {
    S v, du, dt;
    DUAL_HD D2() = default;
    DUAL_HD D2(double x) : v(static_cast<S>(x)), du(0), dt(0) {}
    DUAL_HD D2(S x, S a, S b) : v(x), du(a), dt(b) {}
    DUAL_HD D2 &operator*=(const D2 &o)
    {
        *this = *this * o;
        return *this;
    }
    DUAL_HD friend D2 operator+(const D2 &a, const D2 &b) { return D2(a.v + b.v, a.du + b.du, a.dt + b.dt); }
    DUAL_HD friend D2 operator-(const D2 &a, const D2 &b) { return D2(a.v - b.v, a.du - b.du, a.dt - b.dt); }
    DUAL_HD friend D2 operator-(const D2 &a) { return D2(-a.v, -a.du, -a.dt); }
    DUAL_HD friend D2 operator*(const D2 &a, const D2 &b)
    {
        return D2(a.v * b.v, a.du * b.v + a.v * b.du, a.dt * b.v + a.v * b.dt);
    }
    DUAL_HD friend D2 operator/(const D2 &a, const D2 &b)
    {
        const S inv = arcp(b.v);
        const S q = a.v * inv;
        return D2(q, (a.du - q * b.du) * inv, (a.dt - q * b.dt) * inv);
    }
    // mixed with constants: scale without forming a dual
    DUAL_HD friend D2 operator+(const D2 &a, double c) { return D2(a.v + static_cast<S>(c), a.du, a.dt); }
    DUAL_HD friend D2 operator+(double c, const D2 &a) { return D2(static_cast<S>(c) + a.v, a.du, a.dt); }
    DUAL_HD friend D2 operator-(const D2 &a, double c) { return D2(a.v - static_cast<S>(c), a.du, a.dt); }
    DUAL_HD friend D2 operator-(double c, const D2 &a) { return D2(static_cast<S>(c) - a.v, -a.du, -a.dt); }
    DUAL_HD friend D2 operator*(const D2 &a, double c)
    {
        const S k = static_cast<S>(c);
        return D2(a.v * k, a.du * k, a.dt * k);
    }
    DUAL_HD friend D2 operator*(double c, const D2 &a) { return a * c; }
    DUAL_HD friend D2 operator/(const D2 &a, double c)
    {
        const S k = fm_div(S(1), static_cast<S>(c));
        return D2(a.v * k, a.du * k, a.dt * k);
    }
    DUAL_HD friend D2 operator/(double c, const D2 &b)
    {
        const S inv = arcp(b.v);
        const S q = static_cast<S>(c) * inv;
        return D2(q, -q * b.du * inv, -q * b.dt * inv);
    }
};

} // namespace foort
