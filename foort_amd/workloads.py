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


# C2: adaptive SquareSubdivisionMesh, Kerr a=0.94, inclination 60 deg: a 256x256
# initial grid refined where neighbouring FourColorScreen values differ, 8 levels
# of subdivision (MaxSubdivide 9), at most 4,000,000 pixels; RLogScale on, so
# geodesics take ~10^2 steps and photon-ring orbits up to the 10^5-step time-out.
C2_TEXT = ("C2: SquareSubdivisionMeshV2 on Kerr a=0.94, 60 deg, 30Mx30M screen: 256x256 initial grid, MaxSubdivide 9 "
           "(8 subdivision levels), IterationPixels 1,000,000, MaxPixels 4,000,000, refining on FourColorScreen; "
           "RK4 adaptive step, FP64, RLogScale on, horizon / boundary r=1000M / time-out 1e5 steps, "
           "FourColorScreen + EquatorialPasses + ClosestRadius")


def c2_config(mesh="SquareSubdivisionMeshV2", christoffel="FD", precision="FP64", schedule="OneShot",
              prefix="c2", initial=65536, max_pixels=4000000, iteration=1000000, max_subdivide=9,
              autotune="None", use_graph=False, tail_geodesics=None):
    """The C2 run as a configuration file text for the foort program (tail_geodesics: the GPU
    group's TailGeodesics, None for the default)."""
    tail = "" if tail_geodesics is None else f" TailGeodesics = {int(tail_geodesics)};"
    return f"""
Metric = {{ Name = "Kerr"; a = 0.94; RLogScale = true; }};
Diagnostics = {{
  FourColorScreen = {{ On = true; UseForMesh = true; }};
  EquatorialPasses = {{ On = true; }};
  ClosestRadius = {{ On = true; }};
}};
Terminations = {{
  Horizon = {{ On = true; Epsilon_Horizon = 0.01; }};
  BoundarySphere = {{ On = true; SphereRadius = 1000.0; }};
  TimeOut = {{ On = true; MaxSteps = 100000; }};
}};
ViewScreen = {{
  Position = {{ t = 0.0; r = 1000.0; theta = {math.radians(60.0)!r}; phi = 0.0; }};
  ScreenSize = {{ x = 30.0; y = 30.0; }};
  Mesh = {{ Type = "{mesh}"; InitialPixels = {initial}; MaxPixels = {max_pixels}; IterationPixels = {iteration};
           MaxSubdivide = {max_subdivide}; InitialSubdivisionToFinal = false; }};
}};
Source = {{ Name = "NoSource"; }};
Integrator = {{ Type = "RK4"; StepSize = 0.03; }};
GPU = {{ Christoffel = "{christoffel}"; Precision = "{precision}"; Schedule = "{schedule}";
        Autotune = "{autotune}"; UseGraph = {"true" if use_graph else "false"};{tail} }};
Output = {{ FilePrefix = "{prefix}"; FileExtension = "dat"; TimeStamp = false; ScreenOutputLevel = 2; }};
"""


# C4: render-server request stream.
C4_TEXT = ("C4: 2000 requests from seed {seed}, Poisson arrivals at {rate:g}/s; final image side log-uniform in "
           "[64, 1024]; adaptive SquareSubdivisionMeshV2 of 2-6 iterations (MaxSubdivide = iterations, initial "
           "side (side-1)/2^(it-1)+1, every pixel with differing neighbours subdivided each iteration); Kerr "
           "a ~ U[0, 0.998] (80%), else Johannsen (a ~ U[0, 0.9], alpha13 ~ U[-0.5, 2]), Manko-Novikov "
           "(a ~ U[0, 0.9], alpha3 ~ U[0, 5]) or Rasheed-Larsen (m = 1, a ~ U[0, 0.9], p = 2, q ~ U[2, 4]); "
           "inclination ~ U[5, 90] deg; 30Mx30M screen, RLogScale on, mixed precision, RK4, horizon / boundary "
           "r=1000M / time-out 1e5 steps, FourColorScreen (mesh value) + EquatorialPasses + ClosestRadius")


def c4_requests(seed=2024, n=2000, rate=40.0):
    """The C4 request stream: [(arrival time in s, RunDesc, MeshDesc, info dict)]."""
    import numpy as np
    from . import program as P
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(n):
        t += rng.exponential(1.0 / rate)
        side = int(round(math.exp(rng.uniform(math.log(64), math.log(1024)))))
        iters = int(rng.integers(2, 7))
        init_side = max(2, int(round((side - 1) / 2 ** (iters - 1))) + 1)
        final_side = (init_side - 1) * 2 ** (iters - 1) + 1
        inc = float(rng.uniform(5.0, 90.0))
        kind = "kerr" if rng.uniform() < 0.8 else ["johannsen", "manko_novikov", "rasheed_larsen"][int(rng.integers(3))]
        if kind == "kerr":
            metric, params = A.METRIC_KERR, (float(rng.uniform(0.0, 0.998)),)
        elif kind == "johannsen":
            metric, params = A.METRIC_JOHANNSEN, (float(rng.uniform(0, 0.9)), float(rng.uniform(-0.5, 2.0)), 0, 0, 0)
        elif kind == "manko_novikov":
            metric, params = A.METRIC_MANKONOVIKOV, (float(rng.uniform(0, 0.9)), float(rng.uniform(0, 5.0)))
        else:
            metric, params = A.METRIC_RASHEEDLARSEN, (1.0, float(rng.uniform(0, 0.9)), 2.0, float(rng.uniform(2.0, 4.0)))
        desc = c0(metric=metric, params=params, rlog=True, inclination_deg=inc,
                  christoffel=A.CHRISTOFFEL_ANALYTIC, precision=A.PRECISION_MIXED)
        mesh = P.mesh_desc(P.MESH_SQUARESUBDIVISIONV2, A.DIAGFLAG_FOURCOLORSCREEN, initial_pixels=init_side ** 2,
                           max_pixels=final_side ** 2, iteration_pixels=final_side ** 2, max_subdivide=iters)
        out.append((t, desc, mesh, {"kind": kind, "side": final_side, "iterations": iters}))
    return out
