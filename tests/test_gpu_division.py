"""The integration kernels' binary64 division (foort_amd/csrc/fm_div.hpp: the compiler's
Newton sequence without the v_div_scale_f64 operand scaling) against the compiler's own
a / b on the device and against the host's IEEE division, bit for bit: random operands with
exponents in [-300, 300] (where the scaling is the identity), operands of the path's own
magnitudes, zeros, infinities and NaN in every combination, and operand pairs up to the ends
of the range fm_div.hpp states (|a| >= 2^-968, |b| < 2^1021, exponents at most 767 apart, all
normal), its boundary values included."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _operands(n, lo, hi, rng):
    mant = rng.integers(0, 1 << 52, n, dtype=np.uint64)
    exp = rng.integers(lo + 1023, hi + 1023 + 1, n).astype(np.uint64)
    sign = rng.integers(0, 2, n).astype(np.uint64)
    return ((sign << np.uint64(63)) | (exp << np.uint64(52)) | mant).view(np.float64)


def _same(x, y):
    return (x.view(np.uint64) == y.view(np.uint64)) | (np.isnan(x) & np.isnan(y))


def test_path_division_is_the_compilers_division():
    import foort_amd
    rng = np.random.default_rng(11)
    n = 1 << 22
    a = _operands(n, -300, 300, rng)
    b = _operands(n, -300, 300, rng)
    a[: n // 4] = _operands(n // 4, -40, 40, rng)
    b[: n // 4] = _operands(n // 4, -40, 40, rng)
    special = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1.0, -3.5, 5e-7, 2e6])
    sa, sb = np.meshgrid(special, special)
    a[-sa.size:], b[-sb.size:] = sa.ravel(), sb.ravel()
    qp, qd = np.empty(n), np.empty(n)
    ptr = lambda x: x.ctypes.data_as(C.c_void_p)
    print(f"[division] {n} operand pairs", flush=True)
    assert foort_amd.lib().foort_gpu_check_division(n, ptr(a), ptr(b), ptr(qp), ptr(qd)) == 0
    assert _same(qp, qd).all()
    with np.errstate(all="ignore"):
        host = a / b
    assert _same(qp, host).all()
    # zeros, infinities and NaN behave as IEEE division: signed zeros and infinities
    k = slice(n - sa.size, n)
    assert (np.signbit(qp[k]) == np.signbit(host[k]))[~np.isnan(qp[k])].all()


def _pairs_near_the_limits(n, rng):
    """Operand pairs with exponents ea >= -968, eb <= 1020 and |ea - eb| <= 767 (so the
    quotient is normal), drawn over that whole region, plus its corners."""
    ea = rng.integers(-968, 1021, 4 * n)
    eb = rng.integers(-1022, 1021, 4 * n)
    ok = np.abs(ea - eb) <= 767
    ea, eb = ea[ok][:n], eb[ok][:n]
    corners = [(-968, -968 + 767), (-968, -1022), (-968, -201), (1020, 1020 - 767), (1020, 1020),
               (-200, 567), (567, -200), (253, 1020), (1020, 253), (-700, 67), (67, -700)]
    ea[: len(corners)] = [c[0] for c in corners]
    eb[: len(corners)] = [c[1] for c in corners]
    def make(e):
        mant = rng.integers(0, 1 << 52, e.size, dtype=np.uint64)
        sign = rng.integers(0, 2, e.size).astype(np.uint64)
        return ((sign << np.uint64(63)) | ((e + 1023).astype(np.uint64) << np.uint64(52)) | mant).view(np.float64)
    a, b = make(ea), make(eb)
    # the largest and smallest mantissas at the corners
    a[len(corners):2 * len(corners)] = np.nextafter(np.ldexp(1.0, ea[:len(corners)]), np.inf)
    b[len(corners):2 * len(corners)] = np.nextafter(np.ldexp(2.0, eb[:len(corners)]), 0)
    return a, b


def test_path_division_up_to_the_stated_limits():
    import foort_amd
    rng = np.random.default_rng(12)
    n = 1 << 21
    a, b = _pairs_near_the_limits(n, rng)
    assert (np.abs(a) >= 2.0 ** -968).all() and (np.abs(b) < 2.0 ** 1021).all()
    qp, qd = np.empty(n), np.empty(n)
    ptr = lambda x: x.ctypes.data_as(C.c_void_p)
    print(f"[division] {n} operand pairs up to the stated limits", flush=True)
    assert foort_amd.lib().foort_gpu_check_division(n, ptr(a), ptr(b), ptr(qp), ptr(qd)) == 0
    with np.errstate(all="ignore"):
        host = a / b
    assert np.isfinite(host).all() and (np.abs(host) >= np.finfo(np.float64).tiny).all()
    bad = ~_same(qp, host)
    assert not bad.any(), (int(bad.sum()), a[bad][:4], b[bad][:4])
    assert _same(qd, host).all()
