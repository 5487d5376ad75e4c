"""Accuracy of the portable binary64 functions (foort_amd/csrc/fm_math.h)
against mpmath at 200 bits, on the argument ranges the path uses, plus the
special values.  The same code runs in the GPU kernels; GPU/CPU identity is
covered by the parity tests."""
import ctypes as C
import math

import mpmath as mp
import numpy as np
import pytest

import oracle as O

mp.mp.prec = 200


def _eval(which, x):
    L = O.lib("fm")
    x = np.ascontiguousarray(x, dtype=np.float64)
    a, b = np.zeros_like(x), np.zeros_like(x)
    L.oracle_fm_eval.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.c_void_p, C.c_void_p]
    L.oracle_fm_eval(x.ctypes.data_as(C.c_void_p), len(x), which, a.ctypes.data_as(C.c_void_p),
                     b.ctypes.data_as(C.c_void_p))
    return a, b


def _ulp_err(v, ref):
    ref = mp.mpf(ref)
    if ref == 0:
        return 0.0 if v == 0 else math.inf
    e = mp.floor(mp.log(abs(ref), 2))
    return float(abs(mp.mpf(float(v)) - ref) / mp.power(2, e - 52))


def _max_ulp(vals, xs, f):
    return max(_ulp_err(v, f(mp.mpf(float(x)))) for v, x in zip(vals, xs))


RNG = np.random.default_rng(11)


def test_sincos_theta_range():
    x = RNG.uniform(-0.1, 3.3, 3000)
    s, c = _eval(0, x)
    assert _max_ulp(s, x, mp.sin) < 0.8
    assert _max_ulp(c, x, mp.cos) < 0.8


def test_sincos_wide_range_and_specials():
    x = np.concatenate([RNG.uniform(-1e6, 1e6, 500), [0.0, -0.0, 1e-300, 1.5707963267948966, 3.141592653589793]])
    s, c = _eval(0, x)
    assert _max_ulp(s[:-5], x[:-5], mp.sin) < 1.0
    assert s[-5] == 0.0 and c[-5] == 1.0 and s[-3] == 1e-300
    s, c = _eval(0, np.array([np.inf, -np.inf, np.nan]))
    assert np.isnan(s).all() and np.isnan(c).all()


def test_exp_accuracy_and_limits():
    x = np.concatenate([RNG.uniform(-10, 10, 2000), RNG.uniform(-700, 700, 500)])
    e, _ = _eval(1, x)
    assert _max_ulp(e, x, mp.exp) < 0.55
    e, _ = _eval(1, np.array([0.0, 710.0, -746.0, np.inf, -np.inf, np.nan, 709.7, -740.0]))
    assert e[0] == 1.0 and e[1] == np.inf and e[2] == 0.0 and e[3] == np.inf and e[4] == 0.0
    assert np.isnan(e[5]) and np.isfinite(e[6])
    assert 0 < e[7] < 1e-300  # subnormal range
    assert abs(e[7] - float(mp.exp(-740))) <= 2 ** -1074


@pytest.mark.parametrize("which,n", [(2, 3), (3, 4), (4, 8)])
def test_integer_powers_correctly_rounded(which, n):
    x = np.concatenate([RNG.uniform(0.05, 30.0, 2000), -RNG.uniform(0.05, 5.0, 200)])
    p, _ = _eval(which, x)
    ref = [float(mp.mpf(float(v)) ** n) for v in x]
    assert np.mean(p == np.array(ref)) > 0.999  # double-double: correctly rounded but for rare midpoint cases
    assert _max_ulp(p, x, lambda X: X ** n) < 0.501


def _near(x, y):
    L = O.lib("fm")
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    s, c = np.zeros_like(x), np.zeros_like(x)
    L.oracle_fm_near.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p]
    L.oracle_fm_near(x.ctypes.data_as(C.c_void_p), y.ctypes.data_as(C.c_void_p), len(x),
                     s.ctypes.data_as(C.c_void_p), c.ctypes.data_as(C.c_void_p))
    return s, c


def _ulp(v):
    return math.ulp(abs(float(v))) if v != 0 else 2.0 ** -1074


@pytest.mark.parametrize("h", [1e-7, 3e-6, 2.0 ** -13])
def test_sincos_near_stencil_accuracy(h):
    """sin and cos at theta +- h shifted from theta: within 0.75 ulp, plus an absolute 2^-52 |d|
    where the result is close to zero (the bracket of the addition formula rounds at the size of
    d); across the quadrant boundaries pi/4, 3 pi/4, the equator and near the poles."""
    x = np.concatenate([RNG.uniform(2.0 ** -11, math.pi - 2.0 ** -11, 1500),
                        math.pi / 4 + RNG.uniform(-2 * h, 2 * h, 150), 3 * math.pi / 4 + RNG.uniform(-2 * h, 2 * h, 150),
                        math.pi / 2 + RNG.uniform(-2 * h, 2 * h, 150), [2.0 ** -11, math.pi - 2.0 ** -11]])
    for y in (x + h, x - h):
        d = y - x  # exact
        s, c = _near(x, y)
        for fv, f in ((s, mp.sin), (c, mp.cos)):
            for v, yy, dd in zip(fv, y, d):
                ref = f(mp.mpf(float(yy)))
                err = abs(mp.mpf(float(v)) - ref)
                assert err <= 0.75 * _ulp(float(ref)) + 2.0 ** -52 * abs(dd), (yy, v, float(ref))


def test_sincos_near_agrees_with_direct_and_falls_back():
    """Mostly the same bits as evaluating the point directly; outside its range (|y - x| > 2^-12
    or |x| < 2^-11) it is the direct evaluation; NaN stays NaN."""
    x = RNG.uniform(0.1, 3.0, 4000)
    y = x + 1e-7
    s, c = _near(x, y)
    s2, c2 = _eval(0, y)
    assert np.mean(s == s2) > 0.97 and np.mean(c == c2) > 0.97
    xf = np.array([1e-4, -1e-4, 0.3, 2.0, 1e-300])
    yf = np.array([1e-4 + 1e-7, -1e-4 - 1e-7, 0.3 + 0.01, 2.0 - 2.0 ** -11, 1e-7])
    s, c = _near(xf, yf)
    s2, c2 = _eval(0, yf)
    assert (s == s2).all() and (c == c2).all()
    s, c = _near(np.array([np.nan, 1.0]), np.array([np.nan, np.nan]))
    assert np.isnan(s).all() and np.isnan(c).all()


def test_sincos3_is_direct_at_x_and_near_at_each_stencil_point():
    """fm_sincos3 gives fm_sincos at x and, at each of xp and xm, what fm_sincos_near gives for
    that point alone -- also where one point shifts and the other is evaluated in full (h just
    above 2^-12 with x just below 1, where x + h rounds in the next binade), and at the pole
    threshold |x| = 2^-11."""
    h = 2.0 ** -12 + 2.0 ** -53
    x = np.concatenate([RNG.uniform(0.3, 3.0, 2000), 1.0 - RNG.uniform(0, 2 * h, 2000),
                        RNG.uniform(2.0 ** -12, 2.0 ** -10, 500), [2.0 ** -11, np.nextafter(2.0 ** -11, 0)]])
    for hh in (1e-7, h, 1e-3):
        xp, xm = x + hh, x - hh
        L = O.lib("fm")
        out = np.zeros(6 * len(x))
        L.oracle_fm_stencil.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]
        L.oracle_fm_stencil(x.ctypes.data_as(C.c_void_p), xp.ctypes.data_as(C.c_void_p), xm.ctypes.data_as(C.c_void_p),
                            len(x), out.ctypes.data_as(C.c_void_p))
        out = out.reshape(-1, 6)
        s0, c0 = _eval(0, x)
        sp, cp = _near(x, xp)
        sm, cm = _near(x, xm)
        for got, want in ((out[:, 0], s0), (out[:, 1], c0), (out[:, 2], sp), (out[:, 3], cp), (out[:, 4], sm), (out[:, 5], cm)):
            assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
    # the mixed case occurs: one point within 2^-12 of x, the other not
    dp, dm = np.abs((x + h) - x), np.abs((x - h) - x)
    assert np.any((dp <= 2.0 ** -12) != (dm <= 2.0 ** -12))
