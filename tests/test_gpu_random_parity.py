"""Seeded random configurations, GPU finite-difference path against the oracle bit for bit.

Each case draws a metric and its parameters (the ranges of the render-server stream), RLogScale,
the integrator (RK4, Verlet with and without its velocity iteration), the camera inclination,
screen size and centre, the terminations (ThetaSingularity on or off), the time-out, update
frequencies of every termination and diagnostic (so both the every-step kernels and the gated
ones run), the schedule and its refill and tail options, and a step cap.  The cases are fixed
by their seeds, so a failure names a reproducible configuration.
"""
import math

import numpy as np
import pytest

import oracle as O
from foort_amd import abi as A
from tests.cases import FIELDS, bitwise_equal

pytestmark = pytest.mark.gpu

SEEDS = list(range(48))


def random_case(seed):
    rng = np.random.default_rng(1000 + seed)
    kind = ["kerr", "kerr", "johannsen", "manko_novikov", "rasheed_larsen", "flat"][int(rng.integers(6))]
    if kind == "kerr":
        metric, params = A.METRIC_KERR, (float(rng.uniform(0.0, 0.998)),)
    elif kind == "johannsen":
        metric, params = A.METRIC_JOHANNSEN, (float(rng.uniform(0.0, 0.9)), float(rng.uniform(-0.5, 2.0)), 0.0, 0.0, 0.0)
    elif kind == "manko_novikov":
        metric, params = A.METRIC_MANKONOVIKOV, (float(rng.uniform(0.0, 0.9)), float(rng.uniform(0.0, 5.0)))
    elif kind == "rasheed_larsen":
        metric, params = A.METRIC_RASHEEDLARSEN, (1.0, float(rng.uniform(0.0, 0.9)), 2.0, float(rng.uniform(2.0, 4.0)))
    else:
        metric, params = A.METRIC_FLATSPACE, ()
    verlet = rng.uniform() < 0.3
    terms = A.TERMFLAG_HORIZON | A.TERMFLAG_BOUNDARYSPHERE | A.TERMFLAG_TIMEOUT
    if rng.uniform() < 0.5:
        terms |= A.TERMFLAG_THETASINGULARITY
    if kind == "flat":
        terms &= ~A.TERMFLAG_HORIZON
    rlog = kind == "manko_novikov" or (kind != "flat" and bool(rng.uniform() < 0.6))
    # with RLogScale the steps grow with r; without it a step moves r by about 0.03, so the
    # camera comes closer and the time-out is longer
    r_cam = 1000.0 if rlog else float(rng.uniform(40.0, 400.0))
    if rlog:
        max_steps = int(rng.integers(300, 1500 if kind == "manko_novikov" else 4000))
    else:
        max_steps = int(rng.integers(5000, 30000))
    kw = dict(metric=metric, params=params, rlog=rlog, r_cam=r_cam,
              integrator=A.INTEGRATOR_VERLET if verlet else A.INTEGRATOR_RK4,
              verlet_tol=float(rng.choice([-1.0, 1e-3])),
              theta=float(rng.uniform(math.radians(3.0), math.radians(177.0))),
              screen=(float(rng.uniform(5.0, 40.0)), float(rng.uniform(5.0, 40.0))),
              center=(float(rng.uniform(-3.0, 3.0)), float(rng.uniform(-3.0, 3.0))),
              terms=terms, max_steps=max_steps,
              diags=A.DIAGFLAG_FOURCOLORSCREEN | A.DIAGFLAG_EQUATORIALPASSES | A.DIAGFLAG_CLOSESTRADIUS)
    if kind == "flat":
        kw["screen"] = (float(rng.uniform(100.0, 2000.0)),) * 2
    if rng.uniform() < 0.35:  # gated updates: the kernels that keep counters
        n = lambda: int(rng.choice([0, 1, 2, 3]))  # noqa: E731
        kw.update(horizon_update_n=n(), sphere_update_n=n(), timeout_update_n=n(), theta_update_n=n(),
                  ep_update=A.UpdateFrequency(int(rng.choice([0, 1, 2, 5])), int(rng.integers(2)), int(rng.integers(2))),
                  cr_update=A.UpdateFrequency(int(rng.choice([0, 1, 3])), int(rng.integers(2)), int(rng.integers(2))))
    if rng.uniform() < 0.3:
        kw["step_cap"] = int(rng.integers(50, max_steps))
    sched = {}
    if rng.uniform() < 0.5:
        sched = dict(schedule=A.SCHEDULE_PERSISTENT, refill_lanes=int(rng.choice([1, 16, 64])),
                     refill_interval=int(rng.choice([8, 32, 128])),
                     tail_geodesics=int(rng.choice([0, 1, 1 << 30, 2 ** 32 - 1])))
        if rng.uniform() < 0.3:
            sched["steps_per_launch"] = int(rng.integers(100, 2000))
    side = int(rng.integers(8, 13))
    return kind, kw, sched, side


@pytest.mark.parametrize("seed", SEEDS)
def test_random_configuration_bitexact(seed):
    import foort_amd
    kind, kw, sched, side = random_case(seed)
    print(f"[random] seed={seed} {kind} side={side} {kw} {sched}", flush=True)
    desc = A.make_desc(**kw)
    with foort_amd.Context(A.make_desc(**kw, **sched)) as ctx:
        g = ctx.trace_grid(side)
    o = O.trace("fm", desc, side)
    bad = [k for k in FIELDS if not bitwise_equal(g[k], o[k])]
    assert not bad, (seed, bad)
    assert g["stats"].steps == int(o["steps"].sum())
