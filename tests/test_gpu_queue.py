"""Request queue (foort_gpu_queue_*): requests of one group share one device ring of pixels.

Each request (its own spin, camera inclination and screen side) is run alone through a
context (foort_gpu_trace_pixels, one-shot schedule) and through a shared queue, where
its pixels arrive in several batches interleaved with the other requests' and are
integrated by time-sliced launches (lanes carry unfinished geodesics from launch to
launch) over a ring small enough to wrap.  Every pixel's termination and diagnostic
values are bit for bit those of the lone run, and each slot's step count is the sum of
the lone run's per-pixel steps.
"""
import numpy as np
import pytest

import foort_amd as F
from foort_amd import abi as A
from foort_amd import workloads as W

pytestmark = pytest.mark.gpu

REQUESTS = [(0.1, 17.0, 41), (0.5, 45.0, 64), (0.9, 60.0, 57), (0.998, 85.0, 80)]


def _desc(a, inc, christoffel, precision):
    return W.c0(metric=A.METRIC_KERR, params=(a,), rlog=True, inclination_deg=inc, christoffel=christoffel,
                precision=precision)


def _pixels(rng, side, n):
    flat = rng.choice(side * side, size=n, replace=False)
    return np.stack([flat // side, flat % side], axis=1).astype(np.uint64)


@pytest.mark.parametrize("christoffel,precision,launch_us", [
    (A.CHRISTOFFEL_FD, A.PRECISION_FP64, 400),
    (A.CHRISTOFFEL_ANALYTIC, A.PRECISION_FP64, 0),
    (A.CHRISTOFFEL_ANALYTIC, A.PRECISION_MIXED, 200),
])
def test_queue_matches_lone_requests(christoffel, precision, launch_us):
    print(f"[queue] christoffel={christoffel} precision={precision} launch_us={launch_us}", flush=True)
    rng = np.random.default_rng(11)
    descs = [_desc(a, inc, christoffel, precision) for a, inc, _ in REQUESTS]
    batches = [np.array_split(_pixels(rng, side, 600), 3) for _, _, side in REQUESTS]
    want = []
    for d, (_, _, side), bs in zip(descs, REQUESTS, batches):
        with F.Context(d) as ctx:
            want.append(ctx.trace_pixels(side, np.concatenate(bs)))
    got = [{k: [] for k in ("term", "quadrant", "equat_passes", "closest_radius")} for _ in REQUESTS]
    with F.Queue(descs[0], capacity=1024, max_requests=8) as q:
        slots = [q.add(d, side) for d, (_, _, side) in zip(descs, REQUESTS)]
        nxt = [0] * len(REQUESTS)          # next batch to push per request
        inflight = [None] * len(REQUESTS)  # (first, n, cumulative pixels) of the batch in flight
        pushed = [0] * len(REQUESTS)
        launches = 0
        while any(n < 3 for n in nxt) or any(b is not None for b in inflight):
            for r in range(len(REQUESTS)):
                if inflight[r] is None and nxt[r] < 3:
                    try:
                        first, n = q.push(slots[r], batches[r][nxt[r]])
                    except F.FoortError as e:
                        assert "queue full" in str(e)
                        continue
                    pushed[r] += n
                    inflight[r] = (first, n)
                    nxt[r] += 1
            q.run(launch_us)
            launches += 1
            assert launches < 100000
            done, _, _ = q.progress()
            for r in range(len(REQUESTS)):
                if inflight[r] is not None and int(done[slots[r], 0]) == pushed[r]:
                    out = q.results(*inflight[r])
                    for k in got[r]:
                        got[r][k].append(out[k])
                    inflight[r] = None
        done, total, fin = q.progress()
        assert total == fin == sum(pushed)
        for r in range(len(REQUESTS)):
            for k in got[r]:
                g = np.concatenate(got[r][k])
                if k == "closest_radius":
                    assert np.array_equal(g.view(np.uint64), want[r][k].view(np.uint64)), (r, k)
                else:
                    assert np.array_equal(g, want[r][k]), (r, k)
            assert int(done[slots[r], 1]) == int(want[r]["steps"].astype(np.uint64).sum()), r
            q.remove(slots[r])
    print(f"[queue] {launches} launches", flush=True)
    if launch_us:
        assert launches > 3


def test_queue_rejects_other_groups():
    g = _desc(0.5, 45.0, A.CHRISTOFFEL_ANALYTIC, A.PRECISION_FP64)
    with F.Queue(g, capacity=256, max_requests=2) as q:
        other = _desc(0.5, 45.0, A.CHRISTOFFEL_FD, A.PRECISION_FP64)
        with pytest.raises(F.FoortError, match="group"):
            q.add(other, 32)
        a = q.add(g, 32)
        q.add(_desc(0.7, 30.0, A.CHRISTOFFEL_ANALYTIC, A.PRECISION_FP64), 40)
        with pytest.raises(F.FoortError, match="slot"):
            q.add(g, 32)
        q.remove(a)
        assert q.add(g, 16) == a
        with pytest.raises(F.FoortError, match="queue full"):
            q.push(a, np.zeros((257, 2), np.uint64))
        first, n = q.push(a, np.array([[1, 2], [3, 4]], np.uint64))
        q.run(0)
        out = q.results(first, n)
        with F.Context(g) as ctx:
            ref = ctx.trace_pixels(16, np.array([[1, 2], [3, 4]], np.uint64))
        assert np.array_equal(out["quadrant"], ref["quadrant"])
