"""The C1 image at its full size (Kerr a=0.99, 85 deg, 1024x1024 pixels, time-out 1e5 steps),
checked through properties that do not need the oracle on every pixel:

* the persistent schedule (lanes parked and refilled, the quad-lane tail kernel for the last
  geodesics) gives the one-shot kernel's results bit for bit on all 1,048,576 pixels;
* a spread sample of pixels is the oracle's, bit for bit;
* the mixed-precision persistent path differs from the FP64 finite-difference path in at most
  0.1% of the integer diagnostics, and only on pixels with a neighbour of another value (the
  photon-ring and shadow boundaries).
"""
import numpy as np
import pytest

import oracle as O
from foort_amd import abi as A
from tests.cases import FIELDS, bitwise_equal, c1, sample_pixels

pytestmark = pytest.mark.gpu

SIDE = 1024


@pytest.fixture(scope="module")
def fd_oneshot():
    import foort_amd
    print("[full] C1 1024x1024 FD one-shot", flush=True)
    with foort_amd.Context(c1(schedule=A.SCHEDULE_ONESHOT)) as ctx:
        return ctx.trace_grid(SIDE)


def test_c1_persistent_equals_oneshot_on_every_pixel(fd_oneshot):
    import foort_amd
    print("[full] C1 1024x1024 FD persistent", flush=True)
    with foort_amd.Context(c1(schedule=A.SCHEDULE_PERSISTENT)) as ctx:
        p = ctx.trace_grid(SIDE)
    for k in FIELDS:
        assert bitwise_equal(p[k], fd_oneshot[k]), k
    assert p["stats"].steps == fd_oneshot["stats"].steps == int(fd_oneshot["steps"].sum())
    assert p["stats"].launches == 2  # the persistent kernel, then the tail kernel


def test_c1_sample_is_the_oracle(fd_oneshot):
    px = sample_pixels(SIDE, 40, 11)
    idx = (px[:, 0] * SIDE + px[:, 1]).astype(np.int64)
    print(f"[full] oracle on {len(px)} C1 pixels", flush=True)
    o = O.trace("fm", c1(), SIDE, screen_index=px)
    for k in FIELDS:
        assert bitwise_equal(fd_oneshot[k][idx], o[k]), k


def _boundary(img):
    """Pixels with a 4-neighbour of another value."""
    b = np.zeros(img.shape, bool)
    d = img[1:, :] != img[:-1, :]
    b[1:, :] |= d
    b[:-1, :] |= d
    d = img[:, 1:] != img[:, :-1]
    b[:, 1:] |= d
    b[:, :-1] |= d
    return b


@pytest.mark.parametrize("schedule", [A.SCHEDULE_PERSISTENT, A.SCHEDULE_ONESHOT], ids=["persistent", "oneshot"])
def test_c1_mixed_precision_mismatch_is_small_and_on_boundaries(fd_oneshot, schedule):
    import foort_amd
    print(f"[full] C1 1024x1024 mixed schedule={schedule}", flush=True)
    with foort_amd.Context(c1(christoffel=A.CHRISTOFFEL_ANALYTIC, precision=A.PRECISION_MIXED,
                              schedule=schedule)) as ctx:
        m = ctx.trace_grid(SIDE)
    n = SIDE * SIDE
    off = 0
    for k in ("term", "quadrant", "equat_passes"):
        ref = fd_oneshot[k].reshape(SIDE, SIDE)
        diff = m[k].reshape(SIDE, SIDE) != ref
        assert diff.sum() <= 1e-3 * n, (k, int(diff.sum()))
        off += int((diff & ~_boundary(ref)).sum())
    assert off == 0


def test_c1_mixed_oneshot_and_persistent_bit_identical():
    """The whole 1024x1024 image in mixed precision through the one-shot kernel (flags tested
    afresh on every step) and the persistent kernel (flags hoisted, lanes refilled): every field
    bit for bit the same."""
    import foort_amd
    print("[full] C1 1024x1024 mixed one-shot vs persistent", flush=True)
    out = []
    for schedule in (A.SCHEDULE_ONESHOT, A.SCHEDULE_PERSISTENT):
        with foort_amd.Context(c1(christoffel=A.CHRISTOFFEL_ANALYTIC, precision=A.PRECISION_MIXED,
                                  schedule=schedule)) as ctx:
            out.append(ctx.trace_grid(SIDE))
    a, b = out
    for k in FIELDS:
        assert bitwise_equal(a[k], b[k]), k
