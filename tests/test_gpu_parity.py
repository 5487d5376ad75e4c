"""GPU path vs the CPU oracle (portable-math build), bit for bit, through the C ABI.

The GPU kernels and oracle/liboracle_fm.so evaluate the same reference
arithmetic with the same elementary functions (foort_amd/csrc/fm_math.h), so
every field -- termination, the three diagnostics, final position, velocity,
affine parameter and step count -- must agree exactly.
"""
import sys

import numpy as np
import pytest

import oracle as O
from foort_amd import abi as A
from tests.cases import FIELDS, PARITY_CASES, bitwise_equal, c0, c1, johannsen_cfg

pytestmark = pytest.mark.gpu


def _gpu():
    import foort_amd
    return foort_amd


@pytest.mark.parametrize("name", sorted(PARITY_CASES))
def test_parity_bitexact(name):
    mk, side, pixels = PARITY_CASES[name]
    desc = mk()
    print(f"[parity] case={name} side={side} n={'grid' if pixels is None else len(pixels)}", flush=True)
    F = _gpu()
    with F.Context(desc) as ctx:
        if pixels is None:
            g = ctx.trace_grid(side)
            o = O.trace("fm", desc, side)
        else:
            g = ctx.trace_pixels(side, pixels)
            o = O.trace("fm", desc, side, screen_index=pixels)
    bad = [k for k in FIELDS if not bitwise_equal(g[k], o[k])]
    if bad:
        k = bad[0]
        gv = np.asarray(g[k]).reshape(len(g["term"]), -1)
        ov = np.asarray(o[k]).reshape(len(o["term"]), -1)
        i = int(np.nonzero((gv != ov).any(1))[0][0])
        pytest.fail(f"{name}: fields {bad} differ; first pixel {i}: gpu {gv[i]} oracle {ov[i]}")
    assert int(g["stats"].steps) == int(o["steps"].sum())


def test_camera_matches_oracle():
    desc = c1(rlog=True)
    F = _gpu()
    with F.Context(desc) as ctx:
        V, g0 = ctx.camera()
    Vo, g0o = O.camera("fm", desc)
    assert bitwise_equal(V, Vo) and bitwise_equal(g0, g0o)


def test_grid_offset_and_pixel_list_agree():
    """A grid batch starting at `first` equals the same pixels given as a list."""
    desc = c1(rlog=True, max_steps=5000)
    side, first, count = 20, 37, 50
    F = _gpu()
    with F.Context(desc) as ctx:
        a = ctx.trace_grid(side, first, count)
        idx = np.arange(first, first + count)
        b = ctx.trace_pixels(side, np.stack([idx // side, idx % side], 1))
    for k in FIELDS:
        assert bitwise_equal(a[k], b[k]), k


def test_empty_and_single():
    desc = c1(rlog=True, max_steps=5000)
    F = _gpu()
    with F.Context(desc) as ctx:
        e = ctx.trace_grid(16, 0, 0)
        assert len(e["term"]) == 0 and e["stats"].steps == 0
        s = ctx.trace_pixels(16, [[15, 15]])
    o = O.trace("fm", desc, 16, screen_index=np.array([[15, 15]], np.uint64))
    for k in FIELDS:
        assert bitwise_equal(s[k], o[k]), k


def test_invalid_arguments_raise():
    F = _gpu()
    d = c1()
    d.integrator = 7
    with pytest.raises(F.FoortError):
        F.Context(d)
    with F.Context(c1(rlog=True)) as ctx:
        with pytest.raises(F.FoortError):
            ctx.trace_grid(1)
        with pytest.raises(F.FoortError):
            ctx.trace_grid(4, 10, 10)
        with pytest.raises(F.FoortError):
            ctx.trace_pixels(4, [[4, 0]])


def test_device_entry_point_matches_host():
    desc = c1(rlog=True, max_steps=5000)
    side = 24
    F = _gpu()
    with F.Context(desc) as ctx:
        h = ctx.trace_grid(side)
        n = side * side
        d = F.device_results(n)
        st = ctx.trace_device(side, 0, n, d)
        F.synchronize()
        idx = np.arange(n)
        si = F.DeviceArray.from_host(np.stack([idx // side, idx % side], 1).astype(np.uint64))
        d2 = F.device_results(n)
        ctx.trace_device(side, 0, n, d2, screen_index=si)
    for k in FIELDS:
        assert bitwise_equal(d[k].numpy(), h[k]), k
        assert bitwise_equal(d2[k].numpy(), h[k]), k
    assert st.steps == h["stats"].steps and st.geodesics == n


@pytest.mark.parametrize("budget,interval,launch_us", [(37, 0, 0), (4096, 0, 0), (0, 1, 0), (0, 300, 0),
                                                       (0, 0, 100), (500, 7, 50)])
def test_persistent_schedule_bitexact(budget, interval, launch_us):
    """Work-stealing lanes (256 lanes for 2304 pixels, so every lane refills
    several times) with lane state parked in HBM between launches, launches
    bounded by steps or by time, refills every `interval` steps: results must
    equal the oracle bit for bit."""
    desc = c1(rlog=True, max_steps=10000, schedule=A.SCHEDULE_PERSISTENT, lanes_per_cu=1, steps_per_launch=budget,
              refill_interval=interval, launch_us=launch_us)
    side = 48
    print(f"[persistent] budget={budget} interval={interval} launch_us={launch_us}", flush=True)
    F = _gpu()
    with F.Context(desc) as ctx:
        g = ctx.trace_grid(side)
    o = O.trace("fm", desc, side)
    for k in FIELDS:
        assert bitwise_equal(g[k], o[k]), k
    bounded = budget == 37 or launch_us > 0
    assert g["stats"].launches > (1 if bounded else 0)
    if budget == 0 and launch_us == 0:
        assert g["stats"].launches in (1, 2)  # one launch, and the four-lane tail kernel
    assert g["stats"].steps == o["steps"].sum()


TAIL_CASES = {
    "kerr_rlog_rk4": lambda **k: c1(rlog=True, max_steps=10000, **k),
    "johannsen_verlet": lambda **k: johannsen_cfg(**k),
    "rasheed_larsen": lambda **k: c0(metric=A.METRIC_RASHEEDLARSEN, params=(1.0, 0.5, 2.0, 3.0), rlog=True,
                                     inclination_deg=60.0, max_steps=20000, **k),
    "manko_novikov": lambda **k: c0(metric=A.METRIC_MANKONOVIKOV, params=(0.5, 2.5), rlog=True,
                                    inclination_deg=70.0, max_steps=3000, **k),
}


@pytest.mark.parametrize("tail", [1, 40, 1 << 30, 2 ** 32 - 1])
@pytest.mark.parametrize("name", sorted(TAIL_CASES))
def test_persistent_tail_bitexact(name, tail):
    """The batch's last geodesics are finished by quads of lanes (four lanes per geodesic, the
    four finite-difference metric points evaluated one per lane and exchanged within the quad):
    bit for bit the oracle, whether the tail kernel takes one geodesic, some, nearly all (the
    threshold above the batch: it starts from the geodesics' first steps) or none."""
    mk = TAIL_CASES[name]
    desc = mk(schedule=A.SCHEDULE_PERSISTENT, tail_geodesics=tail)
    side = 40
    print(f"[tail] {name} tail={tail}", flush=True)
    F = _gpu()
    with F.Context(desc) as ctx:
        g = ctx.trace_grid(side)
    o = O.trace("fm", mk(), side)
    for k in FIELDS:
        assert bitwise_equal(g[k], o[k]), k
    assert g["stats"].steps == o["steps"].sum()
    assert g["stats"].launches == (1 if tail == 2 ** 32 - 1 else 2)


# Manko-Novikov a=0.5, alpha3=2.5 at 70 deg on a 40x40 grid: pixel 698 goes through the pole
# (ThetaSingularity off), its velocity turns NaN on step 2344 and its position on step 2345
NAN_CASE = lambda **k: c0(metric=A.METRIC_MANKONOVIKOV, params=(0.5, 2.5), rlog=True,  # noqa: E731
                          inclination_deg=70.0, **k)
NAN_LIMITS = {
    "timeout_3000": dict(max_steps=3000),
    "step_cap_first": dict(max_steps=3000, step_cap=2600),
    "step_cap_only": dict(terms=A.TERMFLAG_HORIZON | A.TERMFLAG_BOUNDARYSPHERE, step_cap=2600),
    "timeout_next_step": dict(max_steps=2345),
    "step_cap_next_step": dict(max_steps=100000, step_cap=2346),
    "gated_counters": dict(max_steps=3000, ep_update=A.UpdateFrequency(2, 0, 0)),
}


@pytest.mark.parametrize("schedule", ["oneshot", "persistent", "persistent_no_tail"])
@pytest.mark.parametrize("limits", sorted(NAN_LIMITS))
def test_nan_geodesic_skips_to_its_last_step(limits, schedule):
    """A geodesic whose state has turned NaN changes nothing but its step counters until
    TimeOut or the step cap ends it; the kernels jump to its last step.  Every field, the step
    count included, is the oracle's (which takes every step), at a time-out or step cap well
    after the NaN, right after it, and with a gated counter (no jump)."""
    kw = dict(NAN_LIMITS[limits])
    sched = {"oneshot": dict(schedule=A.SCHEDULE_ONESHOT),
             "persistent": dict(schedule=A.SCHEDULE_PERSISTENT),
             "persistent_no_tail": dict(schedule=A.SCHEDULE_PERSISTENT, tail_geodesics=2 ** 32 - 1)}[schedule]
    side = 40
    print(f"[nan] {limits} {schedule}", flush=True)
    F = _gpu()
    with F.Context(NAN_CASE(**kw, **sched)) as ctx:
        g = ctx.trace_grid(side)
    o = O.trace("fm", NAN_CASE(**kw), side)
    assert np.isnan(o["final_pos"][698]).all() and np.isnan(o["final_vel"][698]).all()
    for k in FIELDS:
        assert bitwise_equal(g[k], o[k]), k
    assert g["stats"].steps == o["steps"].sum()


def test_persistent_pixel_list_verlet():
    from tests.cases import johannsen_cfg
    desc = johannsen_cfg(
        schedule=A.SCHEDULE_PERSISTENT, lanes_per_cu=1, steps_per_launch=100)
    from tests.cases import sample_pixels
    px = sample_pixels(250, 700, 9)
    F = _gpu()
    with F.Context(desc) as ctx:
        g = ctx.trace_pixels(250, px)
    o = O.trace("fm", desc, 250, screen_index=px)
    for k in FIELDS:
        assert bitwise_equal(g[k], o[k]), k


@pytest.mark.parametrize("schedule", [A.SCHEDULE_ONESHOT, A.SCHEDULE_PERSISTENT])
def test_graph_batches_replay_bitexact(schedule):
    """use_graph: every batch that is one launch runs as a launch of the context's executable
    HIP graph (index upload, kernels, result copies), which each batch's capture updates in
    place -- sizes, grids and arguments change from batch to batch, as an adaptive mesh's do.
    Batches of different sizes, a grid batch and a repeat, with eager work of another context
    between them, all equal the oracle."""
    from tests.cases import sample_pixels
    desc = c1(rlog=True, max_steps=10000, schedule=schedule, use_graph=1)
    eager = c1(rlog=True, max_steps=10000)
    px = sample_pixels(96, 900, 11)
    o = O.trace("fm", desc, 96, screen_index=px)
    print(f"[graph] schedule={schedule}", flush=True)
    F = _gpu()
    with F.Context(desc) as ctx, F.Context(eager) as ectx:
        for rep, n in enumerate((900, 100, 517, 900, 3)):
            g = ctx.trace_pixels(96, px[:n])
            assert g["stats"].graph_launches == 1, rep
            for k in FIELDS:
                assert bitwise_equal(g[k], o[k][:n]), (rep, k)
            assert g["stats"].steps == int(o["steps"][:n].sum())
            e = ectx.trace_pixels(96, px[::-1])
            assert e["stats"].graph_launches == 0
            assert bitwise_equal(e["quadrant"], o["quadrant"][::-1])
        g2 = ctx.trace_grid(40)
        o2 = O.trace("fm", desc, 40)
        assert g2["stats"].graph_launches == 1
        for k in FIELDS:
            assert bitwise_equal(g2[k], o2[k]), k
        g3 = ctx.trace_pixels(96, px[:100])
        for k in FIELDS:
            assert bitwise_equal(g3[k], o[k][:100]), k


def test_autotune_picks_options_and_keeps_results():
    from tests.cases import sample_pixels
    desc = c1(rlog=True, max_steps=10000)
    px = sample_pixels(64, 1500, 5)
    print("[autotune]", flush=True)
    F = _gpu()
    with F.Context(desc) as ctx:
        chosen, st = ctx.autotune(64, screen_index=px, tune_schedule=True)
        assert chosen.schedule in (A.SCHEDULE_ONESHOT, A.SCHEDULE_PERSISTENT)
        if chosen.schedule == A.SCHEDULE_ONESHOT:
            assert chosen.block_size in (64, 128, 256)
        else:
            assert chosen.refill_interval in (8, 32, 128) and chosen.refill_lanes in (1, 16, 64)
        assert st.geodesics == len(px) and st.device_ms > 0
        g = ctx.trace_grid(48)
    o = O.trace("fm", desc, 48)
    for k in FIELDS:
        assert bitwise_equal(g[k], o[k]), k


def _initial_velocities(desc, side):
    """The initial velocities of a side x side grid (FOORT/ViewScreen.cpp:85-110, from the
    oracle's camera frame), to count the warnings the reference prints for them."""
    import math
    V, g0 = O.camera("fm", desc)
    r, th = desc.camera_position[1], desc.camera_position[2]
    out = []
    for row in range(side):
        for col in range(side):
            al = desc.screen_center[0] + desc.screen_size[0] * (row / (side - 1) - 0.5)
            be = desc.screen_center[1] + desc.screen_size[1] * (col / (side - 1) - 0.5)
            d = math.sqrt(r * r + al * al + be * be)
            v = V.T @ np.array([-1.0, -r / d, -be / r / d, al / r / math.sin(th) / d])
            out.append(v / (g0 @ v))
    return np.array(out)


# a camera at r = 1.5 in the ergoregion of Kerr a = 0.99: a third of the initial velocities
# point to the future and outward, which the reference warns about pixel by pixel
ERGO_CAMERA = dict(params=(0.99,), inclination_deg=90.0, r_cam=1.5, screen=(4.0, 4.0), max_steps=2000)


@pytest.mark.parametrize("schedule", [A.SCHEDULE_ONESHOT, A.SCHEDULE_PERSISTENT])
def test_initial_velocity_warnings_and_nonfinite_results_counted(schedule):
    """foort_stats counts the initial velocities the reference warns about
    (FOORT/ViewScreen.cpp:112-116) and the geodesics ending with a non-finite state."""
    side = 8
    print(f"[warn] ergoregion camera, schedule {schedule}", flush=True)
    F = _gpu()
    with F.Context(c0(schedule=schedule, **ERGO_CAMERA)) as ctx:
        g = ctx.trace_grid(side)
    o = O.trace("fm", c0(**ERGO_CAMERA), side)
    for k in FIELDS:
        assert bitwise_equal(g[k], o[k]), k
    v = _initial_velocities(c0(**ERGO_CAMERA), side)
    st = g["stats"]
    assert (st.future_pointing, st.outward_pointing) == (int((v[:, 0] > 0).sum()), int((v[:, 1] > 0).sum()))
    assert st.future_pointing > 0 and st.outward_pointing > 0
    bad = ~np.isfinite(np.concatenate([o["final_pos"], o["final_vel"]], axis=1)).all(axis=1)
    assert st.nonfinite == int(bad.sum())
    # the Manko-Novikov grid with one geodesic through the pole: one non-finite result
    with F.Context(NAN_CASE(max_steps=3000, schedule=schedule)) as ctx:
        n = ctx.trace_grid(40)
    assert (n["stats"].nonfinite, n["stats"].future_pointing, n["stats"].outward_pointing) == (1, 0, 0)


THETA_STENCIL_CASES = {
    # every theta +- h evaluated in full (|d| > 2^-12)
    "derivative_h_1e-3": (lambda: c0(params=(0.7,), rlog=True, inclination_deg=60.0, max_steps=20000,
                                     derivative_h=1e-3), 10),
    # h = 2^-12 + 2^-53: for theta in [1, 2) the rounding of theta +- h is a tie, and |d| lands on
    # 2^-12 or one ulp above it with the last bit of theta -- both choices within fm_sincos3 and
    # from one lane to the next; below 1 every |d| is h, above the threshold
    "derivative_h_near_2^-12": (lambda: c0(params=(0.7,), rlog=True, inclination_deg=60.0, max_steps=20000,
                                           derivative_h=2.0 ** -12 + 2.0 ** -53), 10),
    # camera 3e-4 rad from the pole: theta below 2^-11, the stencil evaluated in full there
    "camera_near_pole": (lambda: c0(params=(0.5,), rlog=True, theta=3e-4, max_steps=20000), 10),
}


@pytest.mark.parametrize("name", sorted(THETA_STENCIL_CASES))
def test_theta_stencil_fallbacks_bitexact(name):
    """sin and cos at theta +- h come from theta's reduced sine and cosine where |d| <= 2^-12 and
    |theta| >= 2^-11, and from a full evaluation elsewhere (fm_sincos3): both branches, and
    lanes taking different ones, bit for bit the oracle's."""
    mk, side = THETA_STENCIL_CASES[name]
    desc = mk()
    print(f"[parity] theta stencil case={name} side={side}", flush=True)
    F = _gpu()
    with F.Context(desc) as ctx:
        g = ctx.trace_grid(side)
    o = O.trace("fm", desc, side)
    bad = [k for k in FIELDS if not bitwise_equal(g[k], o[k])]
    assert not bad, (name, bad)
    assert g["stats"].steps == int(o["steps"].sum())
