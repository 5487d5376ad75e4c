"""Pins of the CPU oracle (oracle/foort_oracle.c).

1. Against the reference's own output files (FOORT/Output/Test_*.dat and
   debug_*.dat, converted by tests/golden/make_golden.py): every integer
   diagnostic of both 250 x 250 runs, with C-library math and with the
   portable functions the GPU uses.
2. Against the reference's own sources compiled into oracle/_ref: final states
   of the parity cases recorded in tests/golden/ref_states.npz, bit for bit,
   and -- when oracle/_ref is built -- metric, Christoffel symbols, single
   integration steps and initial conditions at random points, bit for bit.
3. Against the reference's sample run FOORT/Output/sampleoutput_simplemesh_*
   (an older version with a mirrored screen), on pixels away from boundaries.
"""
import numpy as np
import pytest

import oracle as O
from foort_amd import abi as A
from tests.cases import FIELDS, NO_REFERENCE_COUNTERPART, PARITY_CASES, bitwise_equal, debug_cfg, johannsen_cfg

def _gold(name):
    import os
    return np.load(os.path.join(os.path.dirname(__file__), "golden", name))


@pytest.mark.parametrize("run,kind", [("Test", "libm"), ("Test", "fm"), ("debug", "libm")])
def test_reference_output_files_reproduced(run, kind):
    g = _gold("reference_outputs.npz")
    desc = johannsen_cfg() if run == "Test" else debug_cfg()
    o = O.trace(kind, desc, 250)
    for key, field in (("fcs", "quadrant"), ("ep", "equat_passes")):
        ref = g[f"{run}_{key}"].astype(np.int32)
        mine = o[field].reshape(250, 250)
        assert (mine == ref).all(), f"{run} {key}: {(mine != ref).sum()} of 62500 pixels differ"


@pytest.mark.parametrize("name", sorted(set(PARITY_CASES) - NO_REFERENCE_COUNTERPART))
def test_oracle_matches_reference_states(name):
    g = _gold("ref_states.npz")
    mk, side, px = PARITY_CASES[name]
    o = O.trace("libm", mk(), side) if px is None else O.trace("libm", mk(), side, screen_index=px)
    for k in FIELDS:
        assert bitwise_equal(o[k], g[f"{name}/{k}"]), f"{name}: {k}"


METRIC_DESCS = {
    "kerr": A.make_desc(A.METRIC_KERR, (0.99,)),
    "kerr_rlog": A.make_desc(A.METRIC_KERR, (0.7,), rlog=True),
    "flat": A.make_desc(A.METRIC_FLATSPACE, ()),
    "rasheed_larsen": A.make_desc(A.METRIC_RASHEEDLARSEN, (1.0, 0.6, 2.5, 3.0), rlog=True),
    "johannsen": A.make_desc(A.METRIC_JOHANNSEN, (0.7, 2.0, 0.3, -0.2, 0.4)),
    "manko_novikov": A.make_desc(A.METRIC_MANKONOVIKOV, (0.5, 1.0)),
    "manko_novikov_rlog": A.make_desc(A.METRIC_MANKONOVIKOV, (0.8, -0.5), rlog=True),
}

needs_ref = pytest.mark.skipif(not O.have("ref"), reason="oracle/_ref not built (needs the reference tree)")


def _points(desc, n=40, seed=3):
    rng = np.random.default_rng(seed)
    r = rng.uniform(4.0, 900.0, n)
    u = np.log(r) if desc.rlog_scale else r
    th = rng.uniform(0.05, 3.09, n)
    return [np.array([rng.uniform(-5, 5), u[i], th[i], rng.uniform(0, 6)]) for i in range(n)]


@needs_ref
@pytest.mark.parametrize("mname", sorted(METRIC_DESCS))
def test_metric_and_christoffel_bitexact(mname):
    d = METRIC_DESCS[mname]
    for p in _points(d):
        a, b = O.metric("libm", d, p), O.metric("ref", d, p)
        assert bitwise_equal(a[0], b[0]) and bitwise_equal(a[1], b[1])
        assert bitwise_equal(O.christoffel("libm", d, p), O.christoffel("ref", d, p))


@needs_ref
@pytest.mark.parametrize("mname", sorted(METRIC_DESCS))
@pytest.mark.parametrize("integ,tol", [(A.INTEGRATOR_RK4, 0.001), (A.INTEGRATOR_VERLET, -1.0), (A.INTEGRATOR_VERLET, 1e-4)])
def test_single_step_bitexact(mname, integ, tol):
    d = METRIC_DESCS[mname]
    d.integrator = integ
    d.verlet_tolerance = tol
    rng = np.random.default_rng(5)
    for p in _points(d, 20):
        v = rng.normal(size=4)
        a, b = O.step("libm", d, p, v), O.step("ref", d, p, v)
        assert bitwise_equal(a[0], b[0]) and bitwise_equal(a[1], b[1]) and a[2] == b[2]


@needs_ref
@pytest.mark.parametrize("mname", sorted(METRIC_DESCS))
def test_initial_conditions_bitexact(mname):
    d = METRIC_DESCS[mname]
    d.camera_position[2] = 1.3
    d.screen_size[:] = [14.0, 9.0]
    d.screen_center[:] = [1.5, -0.5]
    for row, col in [(0, 0), (3, 17), (40, 2), (63, 63), (31, 32)]:
        a, b = O.initial("libm", d, 64, row, col), O.initial("ref", d, 64, row, col)
        assert bitwise_equal(a[0], b[0]) and bitwise_equal(a[1], b[1])


def test_oracle_on_the_reference_sample_run():
    """The oracle against FOORT/Output/sampleoutput_simplemesh_* (Kerr a = 0.5, 90 deg, 100 x
    100), written by an older version of the reference whose screen is mirrored (column c of
    the file is column 99 - c here, FourColorScreen's phi halves swapped): 48 pixels away from
    the file's own boundaries agree exactly (the GPU test of the same files covers every pixel)."""
    import math
    g = _gold("sample_outputs.npz")
    img = np.zeros((100, 100), np.int32)
    img[g["simplemesh_rows"].astype(int), g["simplemesh_cols"].astype(int)] = g["simplemesh_FourColorScreen"]
    edge = np.zeros(img.shape, bool)
    for ax in (0, 1):
        d = np.diff(img, axis=ax) != 0
        if ax == 0:
            edge[1:] |= d
            edge[:-1] |= d
        else:
            edge[:, 1:] |= d
            edge[:, :-1] |= d
    rng = np.random.default_rng(17)
    cand = np.argwhere(~edge)
    pick = cand[rng.choice(len(cand), 48, replace=False)]
    desc = A.make_desc(A.METRIC_KERR, (0.5,), rlog=False, theta=math.pi / 2, screen=(15.0, 15.0), max_steps=1000000,
                       horizon_eps=0.01, diags=A.DIAGFLAG_FOURCOLORSCREEN)
    px = np.stack([pick[:, 0], 99 - pick[:, 1]], 1).astype(np.uint64)
    o = O.trace("fm", desc, 100, screen_index=px)
    swap = np.array([0, 2, 1, 4, 3])
    assert (swap[o["quadrant"]] == img[pick[:, 0], pick[:, 1]]).all()
