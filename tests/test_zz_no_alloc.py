"""foort_gpu_trace_device_noalloc: the device-resident batch call that allocates nothing (its
counters and lane slots live in the library's static device array).  Through the C ABI, its
results equal foort_gpu_trace_device's bit for bit on every field, on two calls in a row --
the second on a smaller batch, so anything the first call left in the shared array would
show -- for both schedules (the persistent one with its four-lane tail), grid indices and
pixel lists.  A lanes_per_cu that needs more lanes than the array holds is refused before
anything runs."""
import numpy as np
import pytest

import foort_amd as F
from foort_amd import abi as A
from tests.cases import FIELDS, bitwise_equal, c0, sample_pixels

pytestmark = pytest.mark.gpu

SIDE = 96


def _fetch(res):
    return {k: v.numpy() for k, v in res.items()}


def _same(a, b):
    bad = [k for k in FIELDS if not bitwise_equal(a[k], b[k])]
    assert not bad, f"fields differ: {bad}"


@pytest.mark.parametrize("schedule", [A.SCHEDULE_ONESHOT, A.SCHEDULE_PERSISTENT])
@pytest.mark.parametrize("pixel_list", [False, True])
def test_noalloc_matches_trace_device(schedule, pixel_list):
    desc = c0(params=(0.9,), rlog=True, inclination_deg=70.0, max_steps=20000, schedule=schedule)
    print(f"[noalloc] schedule={schedule} pixel_list={pixel_list}", flush=True)
    with F.Context(desc) as ctx:
        for first, count in ((0, SIDE * SIDE), (517, 1500)):  # the second call is smaller
            si = None
            if pixel_list:
                si = F.DeviceArray.from_host(sample_pixels(SIDE, count, seed=count))
            want_res, got_res = F.device_results(count), F.device_results(count)
            st_want = ctx.trace_device(SIDE, first if si is None else 0, count, want_res, screen_index=si)
            st_got = ctx.trace_device(SIDE, first if si is None else 0, count, got_res, screen_index=si,
                                      noalloc=True)
            _same(_fetch(want_res), _fetch(got_res))
            assert st_got.steps == st_want.steps > 0
            assert st_got.geodesics == st_want.geodesics == count
            assert (st_got.future_pointing, st_got.outward_pointing, st_got.nonfinite) == \
                   (st_want.future_pointing, st_want.outward_pointing, st_want.nonfinite)


def test_noalloc_refuses_more_lanes_than_it_holds():
    # 4096 lanes per CU on every CU is more than the static array's 131072 lanes: refused on
    # the host, before any kernel runs
    desc = c0(schedule=A.SCHEDULE_PERSISTENT, lanes_per_cu=4096)
    with F.Context(desc) as ctx:
        res = F.device_results(64, fields=("term",))
        with pytest.raises(F.FoortError, match="at most 131072"):
            ctx.trace_device(8, 0, 64, res, noalloc=True)
