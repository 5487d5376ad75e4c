"""The analytic right-hand sides (foort_amd/csrc/foort_path.hpp, compiled
for the host into libfoort_path_host.so) against each other and against the
reference's finite differences:

* the hand-derived closed form for Kerr equals the dual-number derivative of
  the reference metric formula to rounding (1e-14 relative);
* both agree with the finite-difference right-hand side (the reference's
  Christoffel symbols, FOORT/Metric.cpp:14-55) within its truncation and
  cancellation error (a few 1e-6 relative at h = 1e-7);
* the binary32 evaluation stays within 1e-5 relative;
* the instantiations with RLogScale fixed at compile time give the same values.
"""
import ctypes as C
import os

import numpy as np
import pytest

from foort_amd import abi as A

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "foort_amd", "libfoort_path_host.so")


def _rhs(L, d, kind, p, v):
    out = np.zeros(4)
    p = np.ascontiguousarray(p, float)
    v = np.ascontiguousarray(v, float)
    L.path_rhs(C.byref(d), kind, p.ctypes.data_as(C.c_void_p), v.ctypes.data_as(C.c_void_p),
               out.ctypes.data_as(C.c_void_p))
    return out


DESCS = {
    "kerr": A.make_desc(A.METRIC_KERR, (0.99,)),
    "kerr_rlog": A.make_desc(A.METRIC_KERR, (0.7,), rlog=True),
    "schwarzschild": A.make_desc(A.METRIC_KERR, (0.0,)),
    "johannsen": A.make_desc(A.METRIC_JOHANNSEN, (0.7, 2.0, 0.3, -0.2, 0.4)),
    "manko_novikov": A.make_desc(A.METRIC_MANKONOVIKOV, (0.5, 1.0)),
    "rasheed_larsen_rlog": A.make_desc(A.METRIC_RASHEEDLARSEN, (1.0, 0.6, 2.5, 3.0), rlog=True),
    "flat": A.make_desc(A.METRIC_FLATSPACE, ()),
}


@pytest.mark.parametrize("name", sorted(DESCS))
def test_analytic_formulations_agree(name):
    L = C.CDLL(LIB)
    d = DESCS[name]
    rng = np.random.default_rng(7)
    for _ in range(150):
        r = rng.uniform(3.0, 500.0)
        u = np.log(r) if d.rlog_scale else r
        p = np.array([0.0, u, rng.uniform(0.1, 3.0), 0.0])
        v = rng.normal(size=4)
        fd, dual, closed, f32, closed_rl, f32_rl = (_rhs(L, d, k, p, v) for k in range(6))
        scale = np.abs(dual).max()
        assert np.abs(closed - dual).max() <= 1e-14 * scale
        assert np.abs(fd - dual).max() <= 5e-6 * scale
        assert np.abs(f32 - dual).max() <= 1e-5 * scale
        # RLogScale fixed at compile time (the kernels' instantiations): the same values
        assert (closed_rl == closed).all() and (f32_rl == f32).all()


@pytest.mark.parametrize("name", sorted(DESCS))
def test_nan_state_gives_nan_derivatives(name):
    """The kernels jump a geodesic whose position and velocity are all NaN to its last step,
    which is exact only if every later step stays all NaN: each right-hand side (finite
    differences, dual numbers, closed form, binary32) maps an all-NaN state to all-NaN."""
    L = C.CDLL(LIB)
    d = DESCS[name]
    nan = np.full(4, np.nan)
    for kind in range(6):
        assert np.isnan(_rhs(L, d, kind, nan, nan)).all(), kind


def test_schwarzschild_radial_infall_known_answer():
    """Radial null geodesic in Schwarzschild: d2r/dlambda2 = 0 for (v^t, v^r) = (1/(1-2/r), -1)
    (E = 1, so dr/dlambda is constant), and d2t/dlambda2 = -2 v^t v^r / (r^2 (1-2/r))."""
    L = C.CDLL(LIB)
    d = DESCS["schwarzschild"]
    for r in (3.0, 10.0, 100.0):
        f = 1 - 2 / r
        v = np.array([1 / f, -1.0, 0.0, 0.0])
        out = _rhs(L, d, 2, [0.0, r, 1.0, 0.0], v)
        assert abs(out[1]) < 1e-15
        assert out[0] == pytest.approx(-2 * v[0] * v[1] / (r * r * f), rel=1e-14)


def test_exp_of_the_analytic_path_within_two_ulp():
    """The table-free binary64 exponential of the analytic formulations (Cody-Waite reduction,
    Taylor polynomial to r^13) against numpy's exp: at most 2 ulp over the arguments the paths
    use (log r of the camera radius and below, the Manko-Novikov potentials' exponents), and
    the special values."""
    L = C.CDLL(LIB)
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(-40.0, 12.0, 200000), rng.uniform(-700.0, 700.0, 20000),
                        np.log(np.array([1.0, 2.0, 1000.0, 1.0 + 1e-12])), [0.0, -0.0, 1e-300, -1e-300]])
    out = np.zeros_like(x)
    L.path_exp(x.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p), C.c_uint64(len(x)))
    ref = np.exp(x)
    ulp = np.abs(out - ref) / np.spacing(ref)
    assert ulp.max() <= 2.0, (ulp.max(), x[ulp.argmax()])
    sp = np.array([800.0, np.inf, -800.0, -np.inf, np.nan])
    o = np.zeros_like(sp)
    L.path_exp(sp.ctypes.data_as(C.c_void_p), o.ctypes.data_as(C.c_void_p), C.c_uint64(len(sp)))
    assert o[0] == np.inf and o[1] == np.inf and o[2] == 0.0 and o[3] == 0.0 and np.isnan(o[4])


def test_stage_sincos_shift_within_a_few_ulp():
    """The analytic RK4 stages take sin and cos of theta by shifting those of the step's start
    (addition formulas with Taylor polynomials of the shift, |d| <= 0.1; the full evaluation
    beyond): within 4 ulp of the larger of the start's and the stage's value, which is the size
    of the rounding of the stage's theta itself (theta0 + d, rounded to binary64)."""
    L = C.CDLL(LIB)
    rng = np.random.default_rng(5)
    n = 200000
    th0 = np.concatenate([rng.uniform(-0.2, np.pi + 0.2, n), rng.uniform(-1e-3, 1e-3, 20000),
                          np.pi + rng.uniform(-1e-3, 1e-3, 20000)])
    d = np.concatenate([rng.uniform(-0.1, 0.1, n), rng.uniform(-2e-3, 2e-3, 40000)])
    d[:2000] = rng.uniform(0.1, 1.0, 2000) * np.sign(rng.uniform(-1, 1, 2000))  # the fallback
    s = np.zeros_like(th0)
    c = np.zeros_like(th0)
    L.path_sincos_shift(th0.ctypes.data_as(C.c_void_p), d.ctypes.data_as(C.c_void_p), s.ctypes.data_as(C.c_void_p),
                        c.ctypes.data_as(C.c_void_p), C.c_uint64(len(th0)))
    th = th0 + d
    for got, ref, start in ((s, np.sin(th), np.sin(th0)), (c, np.cos(th), np.cos(th0))):
        scale = np.spacing(np.maximum(np.abs(ref), np.abs(start)))
        err = np.abs(got - ref) / scale
        assert err.max() <= 4.0, (err.max(), th0[err.argmax()], d[err.argmax()])


@pytest.mark.parametrize("a", [0.0, 0.5, 0.9, 0.998])
def test_closed_form_kerr_near_horizon_and_poles(a):
    """The closed-form Kerr right-hand side (factored t and phi rows, r and theta rows contracted
    directly) against the dual-number derivative of the reference formula where its terms are
    largest: r within 1e-4..1 of the outer horizon and theta within 1e-5..0.1 of a pole."""
    import math
    L = C.CDLL(LIB)
    d = A.make_desc(A.METRIC_KERR, (a,))
    rh = 1.0 + math.sqrt(1.0 - a * a)
    rng = np.random.default_rng(17)
    for _ in range(300):
        r = rh * (1.0 + 10.0 ** rng.uniform(-4, 0)) if rng.random() < 0.5 else rng.uniform(2.0, 1000.0)
        pole = 10.0 ** rng.uniform(-5, -1)
        th = [pole, math.pi - pole, rng.uniform(0.1, 3.0)][int(rng.integers(3))]
        p = np.array([0.0, r, th, 0.0])
        v = rng.normal(size=4)
        dual, closed = _rhs(L, d, 1, p, v), _rhs(L, d, 2, p, v)
        assert np.abs(closed - dual).max() <= 1e-12 * np.abs(dual).max(), (r, th)
