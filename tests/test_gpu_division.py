"""The integration kernels' binary64 division (foort_amd/csrc/fm_div.hpp: the compiler's
Newton sequence without the v_div_scale_f64 operand scaling) against the compiler's own
a / b on the device and against the host's IEEE division, bit for bit: random operands with
exponents in [-300, 300] (where the scaling is the identity), operands of the path's own
magnitudes, and zeros, infinities and NaN in every combination."""
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
