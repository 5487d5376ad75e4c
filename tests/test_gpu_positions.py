"""GeodesicPosition on the GPU (foort_gpu_trace_positions) against the CPU oracle.

The finite-difference path is bit for bit the oracle (portable-math build), so every kept
position -- which saves the UpdateFrequency gating makes, which of them the OutputSteps
decimation keeps, and the coordinates themselves -- is compared exactly, together with the
other results of the batch.  The oracle's saving and decimation are pinned against the
reference program's own GeodesicPosition files in tests/test_host.py.
"""
import os

import numpy as np
import pytest

import oracle as O
import foort_amd as F
from foort_amd import abi as A
from foort_amd import program as P
from foort_amd import workloads as W

pytestmark = pytest.mark.gpu

ALL = A.DIAGFLAG_FOURCOLORSCREEN | A.DIAGFLAG_EQUATORIALPASSES | A.DIAGFLAG_CLOSESTRADIUS


def _desc(integrator=A.INTEGRATOR_RK4, update=(1, 0, 1), keep=0, **kw):
    d = W.c0(params=(0.9,), rlog=True, inclination_deg=75.0, diags=ALL | A.DIAGFLAG_GEODESICPOSITION,
             integrator=integrator, max_steps=5000, **kw)
    d.gp_update = A.UpdateFrequency(*update)
    d.gp_output_steps = keep
    return d


CASES = [
    ("every step, all kept", dict(update=(1, 0, 1), keep=0)),
    ("every step, 25 kept", dict(update=(1, 0, 1), keep=25)),
    ("every 7th, 10 kept", dict(update=(7, 1, 0), keep=10)),
    ("start and finish", dict(update=(0, 1, 1), keep=0)),
    ("finish only, Verlet", dict(integrator=A.INTEGRATOR_VERLET, update=(0, 0, 1), keep=3)),
    ("nothing saved", dict(update=(0, 0, 0), keep=5)),
]


@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
def test_positions_match_oracle(name, kw):
    print(f"[positions] {name}", flush=True)
    d = _desc(**kw)
    side = 24
    rng = np.random.default_rng(5)
    flat = rng.choice(side * side, 200, replace=False)
    si = np.stack([flat // side, flat % side], 1).astype(np.uint64)
    with F.Context(d) as ctx:
        g = ctx.trace_positions(side, si)
    o = O.trace_positions("fm", d, side, screen_index=si)
    assert np.array_equal(g["offsets"], o["offsets"])
    assert np.array_equal(g["points"].view(np.uint64), o["points"].view(np.uint64))
    for k in ("term", "quadrant", "equat_passes"):
        assert np.array_equal(g[k], o[k]), k
    assert np.array_equal(g["steps"].astype(np.uint64), o["steps"])
    assert np.array_equal(g["closest_radius"].view(np.uint64), o["closest_radius"].view(np.uint64))
    assert g["stats"].geodesics == len(si)


def test_positions_in_several_writing_passes(monkeypatch):
    """A device buffer of 1000 points: the writing passes cover chunks of pixels (one pixel
    alone when its points exceed the buffer) and give the same points as one pass."""
    d = _desc(update=(3, 0, 1), keep=0)
    with F.Context(d) as ctx:
        one = ctx.trace_positions(16, count=256)
    monkeypatch.setenv("FOORT_GP_CHUNK_POINTS", "1000")
    with F.Context(d) as ctx:
        many = ctx.trace_positions(16, count=256)
    assert many["stats"].launches > 3
    assert np.array_equal(one["offsets"], many["offsets"])
    assert np.array_equal(one["points"].view(np.uint64), many["points"].view(np.uint64))
    o = O.trace_positions("fm", d, 16, count=256)
    assert np.array_equal(many["points"].view(np.uint64), o["points"].view(np.uint64))


@pytest.mark.parametrize("precision", [A.PRECISION_FP64, A.PRECISION_MIXED])
def test_positions_on_the_analytic_path(precision):
    """Analytic Christoffels: the two passes agree (the call checks it) and the kept positions
    are those of a run in which every step is kept, decimated by the rule."""
    d = _desc(update=(1, 0, 1), keep=0, christoffel=A.CHRISTOFFEL_ANALYTIC, precision=precision)
    with F.Context(d) as ctx:
        full = ctx.trace_positions(12, count=144)
    d.gp_output_steps = 9
    with F.Context(d) as ctx:
        cut = ctx.trace_positions(12, count=144)
    assert np.array_equal(full["steps"], cut["steps"])
    for p in range(144):
        pts = full["points"][full["offsets"][p]:full["offsets"][p + 1]]
        S = len(pts)
        J = S // 9 if 9 < S else 0
        idx = list(range(S)) if not J else list(range(0, S, J))
        if J and (S - 1) % J:
            idx[-1] = S - 1
        assert np.array_equal(cut["points"][cut["offsets"][p]:cut["offsets"][p + 1]], pts[idx]), p


GP_CFG = """
Metric = {{ Name = "Kerr"; a = 0.8; RLogScale = true; }};
Diagnostics = {{ GeodesicPosition = {{ On = true; UpdateFrequency = 20; OutputSteps = 12; UseForMesh = true; }};
                EquatorialPasses = {{ On = true; }}; }};
Terminations = {{ Horizon = {{ On = true; }}; BoundarySphere = {{ On = true; SphereRadius = 1000.0; }};
                 TimeOut = {{ On = true; MaxSteps = 3000; }}; }};
ViewScreen = {{ Position = {{ r = 1000.0; theta = 1.2; }}; ScreenSize = {{ x = 20.0; y = 20.0; }};
               Mesh = {{ Type = "{mesh}"; InitialPixels = 100; MaxPixels = 500; IterationPixels = 30;
                        MaxSubdivide = 3; TotalPixels = 400; }}; }};
Integrator = {{ Type = "RK4"; }};
Output = {{ FilePrefix = "unused"; FileExtension = "dat"; TimeStamp = false; ScreenOutputLevel = 0; }};
"""


def _tree(root):
    out = {}
    for dp, _, fs in os.walk(root):
        for f in fs:
            out[os.path.relpath(os.path.join(dp, f), root)] = open(os.path.join(dp, f), "rb").read()
    return out


def _fm_tracer(desc, side, si):
    if desc.diag_flags & A.DIAGFLAG_GEODESICPOSITION:
        return O.trace_positions("fm", desc, side, screen_index=si)
    return O.trace("fm", desc, side, screen_index=si)


@pytest.mark.parametrize("mesh", ["SimpleSquareMesh", "SquareSubdivisionMeshV2"])
def test_program_with_positions_matches_oracle(mesh, tmp_path):
    """The program with GeodesicPosition (also as the mesh value) writes the same files on the GPU
    as with the CPU oracle as its batch tracer."""
    print(f"[positions] program, {mesh}", flush=True)
    cfg = tmp_path / "gp.cfg"
    cfg.write_text(GP_CFG.format(mesh=mesh))
    P.run(str(cfg), output_prefix=str(tmp_path / "gpu" / "run"))
    P.run(str(cfg), output_prefix=str(tmp_path / "cpu" / "run"), tracer=_fm_tracer)
    gpu, cpu = _tree(tmp_path / "gpu"), _tree(tmp_path / "cpu")
    assert sorted(gpu) == sorted(cpu) == ["run_EquatPasses.dat", "run_GeodesicPosition.dat"]
    for k in gpu:
        assert gpu[k] == cpu[k], k
