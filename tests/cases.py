"""Shared run configurations for the parity tests (small enough for the CPU oracle)."""
import math

import numpy as np

from foort_amd import abi as A

from foort_amd.workloads import ALL_DIAGS, ALL_TERMS, c0, c1  # noqa: F401


def johannsen_cfg(**kw):
    """FOORT/Test.cfg, which produced FOORT/Output/Test_*.dat: Johannsen a=0.7 alpha13=2,
    RLogScale, Verlet with tolerance -1, 5 deg, 15x15 screen, 250^2 pixels."""
    args = dict(metric=A.METRIC_JOHANNSEN, params=(0.7, 2.0, 0.0, 0.0, 0.0), rlog=True,
                integrator=A.INTEGRATOR_VERLET, theta=0.087266462599716478846, screen=(15.0, 15.0),
                terms=ALL_TERMS | A.TERMFLAG_THETASINGULARITY, horizon_eps=0.3, max_steps=10000,
                smallest_step=1e-7, verlet_tol=-1.0,
                diags=A.DIAGFLAG_FOURCOLORSCREEN | A.DIAGFLAG_EQUATORIALPASSES,
                value_diag=A.DIAGFLAG_EQUATORIALPASSES)
    args.update(kw)
    return A.make_desc(**args)


def debug_cfg(**kw):
    """The run described by the header of FOORT/Output/debug_*.dat: Johannsen as Test.cfg,
    RK4, 90 deg, 12x12 screen, horizon eps 1e-3, MaxSteps 1e6, EP threshold 1e-4."""
    args = dict(metric=A.METRIC_JOHANNSEN, params=(0.7, 2.0, 0.0, 0.0, 0.0), rlog=True,
                integrator=A.INTEGRATOR_RK4, theta=1.5707963267948966, screen=(12.0, 12.0),
                terms=ALL_TERMS | A.TERMFLAG_THETASINGULARITY, horizon_eps=0.001, max_steps=1000000,
                smallest_step=1e-7, diags=A.DIAGFLAG_FOURCOLORSCREEN | A.DIAGFLAG_EQUATORIALPASSES,
                value_diag=A.DIAGFLAG_EQUATORIALPASSES, ep_threshold=1e-4)
    args.update(kw)
    return A.make_desc(**args)


def sample_pixels(side, n, seed=0):
    """Deterministic spread of n distinct pixels of a side x side grid (row, column)."""
    rng = np.random.default_rng(seed)
    idx = rng.choice(side * side, size=min(n, side * side), replace=False)
    idx.sort()
    return np.stack([idx // side, idx % side], axis=1).astype(np.uint64)


# name -> (desc factory, side, pixel list or None for the full grid)
PARITY_CASES = {
    "kerr099_rlog_85deg_grid16": (lambda: c1(rlog=True, max_steps=10000), 16, None),
    "c0_kerr05_80deg_sample": (lambda: c0(), 128, sample_pixels(128, 24, 1)),
    "c1_kerr099_85deg_sample": (lambda: c1(), 1024, sample_pixels(1024, 24, 2)),
    "test_cfg_johannsen_verlet": (lambda: johannsen_cfg(), 250, sample_pixels(250, 192, 3)),
    "debug_cfg_johannsen_rk4": (lambda: debug_cfg(), 250, sample_pixels(250, 64, 4)),
    "rasheed_larsen_rlog": (lambda: c0(metric=A.METRIC_RASHEEDLARSEN, params=(1.0, 0.6, 2.5, 3.0), rlog=True,
                                       inclination_deg=60.0, max_steps=20000), 12, None),
    "manko_novikov_rlog": (lambda: c0(metric=A.METRIC_MANKONOVIKOV, params=(0.5, 1.0), rlog=True,
                                      inclination_deg=70.0, max_steps=20000), 12, None),
    "flatspace_theta": (lambda: c0(metric=A.METRIC_FLATSPACE, params=(), inclination_deg=45.0,
                                   terms=A.TERMFLAG_BOUNDARYSPHERE | A.TERMFLAG_TIMEOUT | A.TERMFLAG_THETASINGULARITY,
                                   screen=(2000.0, 2000.0)), 6, None),
    "kerr_verlet_tolerance": (lambda: c0(params=(0.9,), rlog=True, integrator=A.INTEGRATOR_VERLET,
                                         inclination_deg=30.0, verlet_tol=1e-3, max_steps=20000), 12, None),
    "update_frequencies": (lambda: c0(params=(0.7,), rlog=True, inclination_deg=50.0, max_steps=3000,
                                      horizon_update_n=2, sphere_update_n=3, timeout_update_n=2,
                                      ep_update=A.UpdateFrequency(0, 1, 1), cr_update=A.UpdateFrequency(5, 0, 0),
                                      terms=ALL_TERMS | A.TERMFLAG_THETASINGULARITY, theta_update_n=4), 12, None),
    "ep_cr_start_finish_only": (lambda: c0(params=(0.3,), rlog=True, ep_update=A.UpdateFrequency(0, 0, 1),
                                           cr_update=A.UpdateFrequency(0, 1, 0), max_steps=5000), 8, None),
    "one_step_timeout": (lambda: c1(max_steps=0), 32, None),
    "step_cap": (lambda: c1(max_steps=100000, step_cap=777), 8, None),
    "off_centre_screen": (lambda: c0(params=(0.95,), rlog=True, screen=(12.0, 8.0), center=(3.0, -2.0),
                                     inclination_deg=17.0, max_steps=8000), 10, None),
}

# cases that exercise a product-only option (no counterpart in the reference)
NO_REFERENCE_COUNTERPART = {"step_cap"}

FIELDS = ("term", "quadrant", "equat_passes", "closest_radius", "final_pos", "final_vel", "lambda", "steps")


def bitwise_equal(a, b):
    """Equal bit for bit; two NaNs count as equal whatever their bits (an invalid operation
    gives a NaN whose sign and payload IEEE 754 leaves open: x86 makes 0xfff8..., the GPU
    0x7ff8..., so a geodesic that turns NaN -- past a pole with ThetaSingularity off, say --
    does so on both sides with different bits)."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    if a.dtype.kind == "f":
        same = a.view(np.uint64) == b.view(np.uint64)
        return a.shape == b.shape and bool(np.all(same | (np.isnan(a) & np.isnan(b))))
    return np.array_equal(a.astype(np.int64), b.astype(np.int64))
