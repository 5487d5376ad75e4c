// fm_div.hpp -- correctly rounded division for the scalar types of the path.
#pragma once

#if defined(__HIPCC__)
#define FD_HD __host__ __device__ __attribute__((always_inline)) inline
#else
#define FD_HD inline
#endif

namespace foort {

// a / b, correctly rounded.  On gfx950 hipcc divides binary64 by v_div_scale_f64 (twice),
// v_rcp_f64, two Newton refinements of the reciprocal, the quotient and one fused
// correction, v_div_fmas_f64 and v_div_fixup_f64 (11 instructions).  The two scalings and
// the scaled multiply-add are the identity (factor 1, plain fma) unless the denominator or the
// quotient is subnormal, |b| >= 2^1021, |a| < 2^-969, or the exponents of a and b differ by
// 768 or more; zero, infinite and NaN operands are settled by v_div_fixup_f64 from the
// operands alone.  This is the same sequence without the scalings (9 instructions), so it
// returns a / b bit for bit for every operand pair outside those ranges.  The path's divisions
// are of metric quantities and step lengths, 30 orders of magnitude inside them.
FD_HD double fm_div(double a, double b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-b, r, 1.0);
    r = __builtin_fma(r, e, r);
    const double q = a * r;
    const double rem = __builtin_fma(-b, q, a);
    return __builtin_amdgcn_div_fixup(__builtin_fma(rem, r, q), b, a);
#else
    return a / b;
#endif
}
FD_HD float fm_div(float a, float b) { return a / b; }
// other scalar types (dual numbers, counting scalars): their own division operators
template <class T>
FD_HD T fm_div(const T &a, const T &b) { return a / b; }
template <class T>
FD_HD T fm_div(double a, const T &b) { return a / b; }
template <class T>
FD_HD T fm_div(const T &a, double b) { return a / b; }

} // namespace foort
