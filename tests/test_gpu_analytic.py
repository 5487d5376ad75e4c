"""Analytic Christoffel symbols (exact derivatives of the reference metrics)
and the mixed binary32/binary64 path, against the finite-difference oracle.

These formulations are not the reference's arithmetic, so agreement is
statistical: integer diagnostics may differ only on a small fraction of
pixels (those next to a diagnostic boundary), and final positions agree to a
stated tolerance on the pixels whose integer diagnostics agree.
"""
import numpy as np
import pytest

import oracle as O
from foort_amd import abi as A
from tests.cases import c0, c1, johannsen_cfg

pytestmark = pytest.mark.gpu


def _gpu():
    import foort_amd
    return foort_amd


def _compare(g, o):
    same = (g["term"] == o["term"]) & (g["quadrant"] == o["quadrant"]) & (g["equat_passes"] == o["equat_passes"])
    rel = np.abs(g["final_pos"] - o["final_pos"]) / np.maximum(np.abs(o["final_pos"]), 1.0)
    return same, rel


CASES = {
    "kerr099_rlog": (lambda **k: c1(rlog=True, max_steps=10000, **k), 32),
    "test_cfg": (lambda **k: johannsen_cfg(**k), 40),
    "rasheed_larsen": (lambda **k: c0(metric=A.METRIC_RASHEEDLARSEN, params=(1.0, 0.6, 2.5, 3.0), rlog=True,
                                      inclination_deg=60.0, max_steps=20000, **k), 24),
    # p = 2m: the constant (p^2 - 4m^2) is exactly zero, and a multiply-add must not make it negative
    "rasheed_larsen_p2m": (lambda **k: c0(metric=A.METRIC_RASHEEDLARSEN, params=(1.0, 0.5, 2.0, 3.0), rlog=True,
                                          inclination_deg=60.0, max_steps=20000, **k), 24),
    "manko_novikov": (lambda **k: c0(metric=A.METRIC_MANKONOVIKOV, params=(0.5, 1.0), rlog=True,
                                     inclination_deg=70.0, max_steps=20000, **k), 24),
    "kerr_verlet_tolerance": (lambda **k: c0(params=(0.9,), rlog=True, integrator=A.INTEGRATOR_VERLET,
                                             inclination_deg=30.0, verlet_tol=1e-3, max_steps=20000, **k), 24),
    "flat": (lambda **k: c0(metric=A.METRIC_FLATSPACE, params=(), inclination_deg=45.0,
                            terms=A.TERMFLAG_BOUNDARYSPHERE | A.TERMFLAG_TIMEOUT, screen=(2000.0, 2000.0),
                            max_steps=100000, **k), 8),
}


@pytest.mark.parametrize("precision", [A.PRECISION_FP64, A.PRECISION_MIXED])
@pytest.mark.parametrize("name", sorted(CASES))
def test_analytic_close_to_fd_oracle(name, precision):
    mk, side = CASES[name]
    kw = dict(christoffel=A.CHRISTOFFEL_ANALYTIC, precision=precision)
    F = _gpu()
    with F.Context(mk(**kw)) as ctx:
        g = ctx.trace_grid(side)
    o = O.trace("fm", mk(), side)
    assert np.isfinite(g["final_pos"]).all()
    same, rel = _compare(g, o)
    frac = 1 - same.mean()
    print(f"[analytic] {name} precision={precision}: integer mismatch {frac:.4f}, "
          f"median rel pos {np.median(rel[same]):.2e}, max {rel[same].max():.2e}", flush=True)
    assert frac <= 0.05
    assert np.median(rel[same]) < (1e-6 if precision == A.PRECISION_FP64 else 1e-4)


@pytest.mark.parametrize("precision", [A.PRECISION_FP64, A.PRECISION_MIXED])
@pytest.mark.parametrize("name", ["kerr099_rlog", "test_cfg", "manko_novikov"])
def test_analytic_persistent_equals_oneshot(name, precision):
    """Bit for bit between the schedules.  The one-shot analytic kernels leave their step loop
    lane by lane, the persistent ones park lanes between launches: a lane stepping past its
    termination (the code-generation issue that term_after_run guards against) would differ."""
    mk, side = CASES[name]
    F = _gpu()
    kw = dict(christoffel=A.CHRISTOFFEL_ANALYTIC, precision=precision)
    with F.Context(mk(**kw)) as ctx:
        a = ctx.trace_grid(side)
    with F.Context(mk(schedule=A.SCHEDULE_PERSISTENT, lanes_per_cu=1, steps_per_launch=50, **kw)) as ctx:
        b = ctx.trace_grid(side)
    for k in ("term", "quadrant", "equat_passes", "steps"):
        assert np.array_equal(a[k], b[k]), k
    for k in ("final_pos", "final_vel", "lambda", "closest_radius"):
        assert np.array_equal(a[k].view(np.uint64), b[k].view(np.uint64)), k


@pytest.mark.parametrize("opts", [dict(lanes_per_cu=4, steps_per_launch=50),
                                  dict(lanes_per_cu=8, refill_lanes=64, refill_interval=8),
                                  dict(lanes_per_cu=4, refill_interval=5, launch_us=300)],
                         ids=["budget50", "refill64_interval8", "time_slices"])
@pytest.mark.parametrize("name,side", [("kerr099_rlog", 64), ("test_cfg", 40)])
def test_mixed_persistent_schedules_equal_oneshot(name, side, opts):
    """The mixed-precision persistent kernel over several workgroups with refills (1024 or 2048
    resident lanes for 1600 or 4096 pixels), launch budgets, refill settings and time slices:
    every result bit for bit the one-shot kernel's."""
    mk, _ = CASES[name]
    F = _gpu()
    kw = dict(christoffel=A.CHRISTOFFEL_ANALYTIC, precision=A.PRECISION_MIXED)
    with F.Context(mk(**kw)) as ctx:
        a = ctx.trace_grid(side)
    print(f"[analytic] mixed persistent {name} {opts}", flush=True)
    with F.Context(mk(schedule=A.SCHEDULE_PERSISTENT, **opts, **kw)) as ctx:
        b = ctx.trace_grid(side)
    for k in ("term", "quadrant", "equat_passes", "steps"):
        assert np.array_equal(a[k], b[k]), k
    for k in ("final_pos", "final_vel", "lambda", "closest_radius"):
        assert np.array_equal(a[k].view(np.uint64), b[k].view(np.uint64)), k


@pytest.mark.parametrize("path", ["fd", "analytic", "mixed"])
def test_kerr_constants_of_motion_on_every_path(path):
    """E = -p_t, L = p_phi, Carter's Q and the null norm at the end of escaping Kerr geodesics
    equal their initial values to each path's accuracy: the finite-difference path as the
    oracle (it is the oracle bit for bit), the closed-form FP64 path better, and the mixed
    path within the rounding of binary32 far-field steps."""
    import math
    import foort_amd
    import oracle as O
    a, side = 0.9, 64
    kw = {"fd": {}, "analytic": dict(christoffel=A.CHRISTOFFEL_ANALYTIC),
          "mixed": dict(christoffel=A.CHRISTOFFEL_ANALYTIC, precision=A.PRECISION_MIXED)}[path]
    d = A.make_desc(A.METRIC_KERR, (a,), rlog=False, inclination_deg=60.0, screen=(20.0, 20.0), max_steps=200000, **kw)
    with foort_amd.Context(d) as ctx:
        g = ctx.trace_grid(side)
    ref = A.make_desc(A.METRIC_KERR, (a,), rlog=False, inclination_deg=60.0, screen=(20.0, 20.0))

    def invariants(p, v):
        m = O.metric("fm", ref, p)[0]
        E, L, pth = -(m[0] @ v), m[3] @ v, m[2, 2] * v[2]
        Q = pth ** 2 + math.cos(p[2]) ** 2 * (-a * a * E * E + L * L / math.sin(p[2]) ** 2)
        return np.array([E, L, Q]), (v @ m @ v) / np.abs(np.outer(v, v) * m).max()

    worst = np.zeros(4)
    for i in np.nonzero(g["term"] == 2)[0][::16]:
        c0, _ = invariants(*O.initial("fm", ref, side, int(i // side), int(i % side)))
        c1, n1 = invariants(g["final_pos"][i], g["final_vel"][i])
        rel = np.abs(c1 - c0) / np.maximum(np.abs(c0), 1e-3)
        worst = np.maximum(worst, np.append(rel, abs(n1)))
    print(f"[invariants] {path}: dE {worst[0]:.2e} dL {worst[1]:.2e} dQ {worst[2]:.2e} null {worst[3]:.2e}", flush=True)
    assert (worst < {"fd": 1e-5, "analytic": 1e-5, "mixed": 1e-4}[path]).all()
