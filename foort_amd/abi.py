"""ctypes mirror of include/foort_gpu.h (the C ABI of the GPU integrator).

The structures below must match the header field for field; the library
reports sizeof(foort_run_desc) through foort_desc_size() and load() checks it.
"""
import ctypes as C
import math

# metrics, FOORT/Metric.h:78-189
METRIC_KERR, METRIC_FLATSPACE, METRIC_RASHEEDLARSEN, METRIC_JOHANNSEN, METRIC_MANKONOVIKOV = range(5)
INTEGRATOR_RK4, INTEGRATOR_VERLET = 0, 1
CHRISTOFFEL_FD, CHRISTOFFEL_ANALYTIC = 0, 1
PRECISION_FP64, PRECISION_MIXED = 0, 1
SCHEDULE_ONESHOT, SCHEDULE_PERSISTENT = 0, 1
# TermBitflag, FOORT/Terminations.h:29-33
TERMFLAG_BOUNDARYSPHERE, TERMFLAG_TIMEOUT, TERMFLAG_HORIZON, TERMFLAG_THETASINGULARITY = 1, 2, 4, 8
# Term, FOORT/Terminations.h:48-67
TERM_CONTINUE, TERM_HORIZON, TERM_BOUNDARYSPHERE, TERM_TIMEOUT, TERM_THETASINGULARITY, TERM_STEPCAP = range(6)
# DiagBitflag, FOORT/Diagnostics.h:33-37
DIAGFLAG_GEODESICPOSITION, DIAGFLAG_FOURCOLORSCREEN, DIAGFLAG_EQUATORIALPASSES, DIAGFLAG_CLOSESTRADIUS = 1, 2, 4, 8

PI_REF = 3.1415926535  # FOORT/Geometry.h:44


class UpdateFrequency(C.Structure):
    _fields_ = [("update_n_steps", C.c_uint64), ("update_start", C.c_int32), ("update_finish", C.c_int32)]


class RunDesc(C.Structure):
    _fields_ = [
        ("metric", C.c_int32), ("rlog_scale", C.c_int32), ("metric_params", C.c_double * 6),
        ("integrator", C.c_int32), ("christoffel", C.c_int32),
        ("step_size", C.c_double), ("smallest_step", C.c_double), ("derivative_h", C.c_double),
        ("verlet_tolerance", C.c_double),
        ("term_flags", C.c_uint32), ("precision", C.c_uint32),
        ("horizon_epsilon", C.c_double), ("horizon_update_n", C.c_uint64),
        ("sphere_radius", C.c_double), ("sphere_update_n", C.c_uint64),
        ("timeout_max_steps", C.c_uint64), ("timeout_update_n", C.c_uint64),
        ("theta_epsilon", C.c_double), ("theta_update_n", C.c_uint64),
        ("diag_flags", C.c_uint32), ("value_diag", C.c_uint32),
        ("ep_threshold", C.c_double), ("ep_update", UpdateFrequency), ("cr_update", UpdateFrequency),
        ("camera_position", C.c_double * 4), ("screen_size", C.c_double * 2), ("screen_center", C.c_double * 2),
        ("step_cap", C.c_uint64), ("schedule", C.c_int32), ("block_size", C.c_int32),
        ("steps_per_launch", C.c_uint32), ("lanes_per_cu", C.c_uint32),
        ("refill_lanes", C.c_uint32), ("refill_interval", C.c_uint32),
        ("mixed_switch_radius", C.c_double), ("mixed_pole_margin", C.c_double),
        ("launch_us", C.c_uint32), ("use_graph", C.c_uint32),
        ("gp_update", UpdateFrequency), ("gp_output_steps", C.c_uint64),
        ("tail_geodesics", C.c_uint32), ("reserved", C.c_uint32),
    ]


class GeodesicResults(C.Structure):
    _fields_ = [
        ("term", C.c_void_p), ("quadrant", C.c_void_p), ("equat_passes", C.c_void_p),
        ("closest_radius", C.c_void_p), ("final_pos", C.c_void_p), ("final_vel", C.c_void_p),
        ("lambda_", C.c_void_p), ("steps", C.c_void_p),
    ]


class Positions(C.Structure):
    """foort_positions (include/foort_gpu.h)"""
    _fields_ = [("offsets", C.POINTER(C.c_uint64)), ("points", C.POINTER(C.c_double)), ("total", C.c_uint64)]


class Stats(C.Structure):
    _fields_ = [("geodesics", C.c_uint64), ("steps", C.c_uint64), ("launches", C.c_uint64),
                ("device_ms", C.c_double), ("future_pointing", C.c_uint64), ("outward_pointing", C.c_uint64),
                ("nonfinite", C.c_uint64), ("graph_launches", C.c_uint64)]


def default_desc():
    """The reference defaults (FOORT/Config.cpp:123-1155, FOORT/Integrators.h:31-57)."""
    d = RunDesc()
    d.metric = METRIC_KERR
    d.rlog_scale = 0
    d.metric_params[0] = 0.5
    d.integrator = INTEGRATOR_RK4
    d.christoffel = CHRISTOFFEL_FD
    d.step_size = 0.03
    d.smallest_step = 1e-12
    d.derivative_h = 1e-7
    d.verlet_tolerance = 0.001
    d.term_flags = TERMFLAG_BOUNDARYSPHERE | TERMFLAG_TIMEOUT
    d.precision = PRECISION_FP64
    d.horizon_epsilon = 0.01
    d.horizon_update_n = 1
    d.sphere_radius = 1000.0
    d.sphere_update_n = 1
    d.timeout_max_steps = 10000
    d.timeout_update_n = 1
    d.theta_epsilon = 1e-5
    d.theta_update_n = 1
    d.diag_flags = DIAGFLAG_FOURCOLORSCREEN
    d.value_diag = DIAGFLAG_FOURCOLORSCREEN
    d.ep_threshold = 0.01
    d.ep_update = UpdateFrequency(1, 1, 1)
    d.cr_update = UpdateFrequency(1, 1, 1)
    d.camera_position[:] = [0.0, 1000.0, PI_REF / 2.0, 0.0]
    d.screen_size[:] = [10.0, 10.0]
    d.screen_center[:] = [0.0, 0.0]
    d.step_cap = 0
    d.schedule = SCHEDULE_ONESHOT
    d.mixed_switch_radius = 10.0
    d.mixed_pole_margin = 0.1
    d.gp_update = UpdateFrequency(1, 0, 1)
    return d


def make_desc(metric=METRIC_KERR, params=(0.5,), rlog=False, integrator=INTEGRATOR_RK4,
              inclination_deg=None, theta=None, r_cam=1000.0, screen=(30.0, 30.0), center=(0.0, 0.0),
              terms=TERMFLAG_HORIZON | TERMFLAG_BOUNDARYSPHERE | TERMFLAG_TIMEOUT,
              diags=DIAGFLAG_FOURCOLORSCREEN | DIAGFLAG_EQUATORIALPASSES | DIAGFLAG_CLOSESTRADIUS,
              value_diag=DIAGFLAG_FOURCOLORSCREEN, max_steps=100000, sphere_radius=1000.0,
              horizon_eps=0.01, step_size=0.03, smallest_step=1e-12, verlet_tol=0.001, **extra):
    d = default_desc()
    d.metric = metric
    d.rlog_scale = 1 if rlog else 0
    for i in range(6):
        d.metric_params[i] = params[i] if i < len(params) else 0.0
    d.integrator = integrator
    if theta is None:
        theta = (inclination_deg if inclination_deg is not None else 90.0) * math.pi / 180.0
    d.camera_position[:] = [0.0, r_cam, theta, 0.0]
    d.screen_size[:] = list(screen)
    d.screen_center[:] = list(center)
    d.term_flags = terms
    d.diag_flags = diags
    d.value_diag = value_diag
    d.timeout_max_steps = max_steps
    d.sphere_radius = sphere_radius
    d.horizon_epsilon = horizon_eps
    d.step_size = step_size
    d.smallest_step = smallest_step
    d.verlet_tolerance = verlet_tol
    for k, v in extra.items():
        setattr(d, k, v)
    return d
