"""The named workloads of the benchmark and the parity tests.

C0  Kerr a=0.5, observer r=1000M, inclination 80 deg, 128x128 SimpleSquareMesh
    over a 30M x 30M screen; RK4 with the adaptive step (epsilon 0.03), finite-
    difference Christoffels (h=1e-7), FP64; horizon (eps 0.01), boundary sphere
    r=1000M, time-out 1e5 steps; FourColorScreen + EquatorialPasses +
    ClosestRadius (value diagnostic FourColorScreen).  RLogScale off (the
    reference default, FOORT/Config.cpp:171-172).
C1  as C0 with Kerr a=0.99, inclination 85 deg, 1024x1024 pixels.
"""
import math

from . import abi as A

ALL_TERMS = A.TERMFLAG_HORIZON | A.TERMFLAG_BOUNDARYSPHERE | A.TERMFLAG_TIMEOUT
ALL_DIAGS = A.DIAGFLAG_FOURCOLORSCREEN | A.DIAGFLAG_EQUATORIALPASSES | A.DIAGFLAG_CLOSESTRADIUS


def c0(**kw):
    args = dict(metric=A.METRIC_KERR, params=(0.5,), rlog=False, inclination_deg=80.0, screen=(30.0, 30.0),
                terms=ALL_TERMS, diags=ALL_DIAGS, max_steps=100000)
    args.update(kw)
    return A.make_desc(**args)


C0_SIDE = 128


def c1(**kw):
    args = dict(params=(0.99,), inclination_deg=85.0)
    args.update(kw)
    return c0(**args)


C1_SIDE = 1024

C1_TEXT = ("Kerr a=0.99, observer r=1000M, inclination 85 deg, 1024x1024 SimpleSquareMesh over a 30Mx30M "
           "screen, RK4 adaptive step (eps 0.03), FP64, horizon / boundary r=1000M / time-out 1e5 steps, "
           "FourColorScreen + EquatorialPasses + ClosestRadius, RLogScale off")
