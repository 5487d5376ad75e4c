"""The host side of the drop-in (libfoort_host.so, include/foort_host.h) on the CPU.

Pins:
1. The configuration reader: the syntax of the reference's configuration files
   and the typing rules of its lookups (a float option given as an integer is
   ignored, and so on), checked against what FOORT/Config.cpp does.
2. The first line of the output files: for the settings of the reference's
   shipped run, byte-identical to the first line of FOORT/Output/Test_*.dat
   (tests/golden/reference_outputs.npz); for many other settings, identical to
   Utilities::GetFirstLineInfoString of the reference sources (oracle/_ref).
3. The meshes: the queue of screen indices of every iteration, element for
   element, against the reference's SimpleSquareMesh, SquareSubdivisionMesh and
   SquareSubdivisionMeshV2 (oracle/_ref) fed the same values.
4. The output handler: the files written for a sequence of batches, byte for
   byte, against the reference's GeodesicOutputHandler (oracle/_ref).
5. The whole program: the driver with the CPU oracle as its batch tracer writes
   the same files as the reference's program loop (oracle/_ref) for adaptive
   meshes, bit-exact arithmetic on both sides.
"""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import oracle as O
from foort_amd import abi as A
from foort_amd import program as P
from foort_amd import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
need_ref = pytest.mark.skipif(not O.have("ref"), reason="oracle/_ref (the reference build) is not built")

FCS, EP, CR = A.DIAGFLAG_FOURCOLORSCREEN, A.DIAGFLAG_EQUATORIALPASSES, A.DIAGFLAG_CLOSESTRADIUS


def _cfg(text):
    return P.Config(text=text)


# ---------------------------------------------------------------------------
# 1. configuration reader
# ---------------------------------------------------------------------------

def test_gpu_group_sets_the_execution_options():
    d = _cfg("""GPU = { Christoffel = "Analytic"; Precision = "Mixed"; Schedule = "Persistent"; StepCap = 5000;
                        BlockSize = 128; StepsPerLaunch = 64; LanesPerCU = 512; RefillLanes = 16;
                        RefillInterval = 8; MixedSwitchRadius = 20.0; MixedPoleMargin = 0.2;
                        LaunchMicroseconds = 300; UseGraph = true; };""").run_desc()
    assert (d.christoffel, d.precision, d.schedule) == (A.CHRISTOFFEL_ANALYTIC, A.PRECISION_MIXED,
                                                        A.SCHEDULE_PERSISTENT)
    assert (d.step_cap, d.block_size, d.steps_per_launch, d.lanes_per_cu) == (5000, 128, 64, 512)
    assert (d.refill_lanes, d.refill_interval, d.launch_us, d.use_graph) == (16, 8, 300, 1)
    assert (d.mixed_switch_radius, d.mixed_pole_margin) == (20.0, 0.2)
    # unknown values fall back to the reference arithmetic and the one-shot schedule
    d = _cfg('GPU = { Christoffel = "Exact"; Precision = "half"; Schedule = "x"; };').run_desc()
    assert (d.christoffel, d.precision, d.schedule) == (A.CHRISTOFFEL_FD, A.PRECISION_FP64, A.SCHEDULE_ONESHOT)


def test_defaults_without_any_setting():
    d = _cfg("").run_desc()
    # Kerr a = 0.5, FourColorScreen, BoundarySphere(1000) + TimeOut(10000), camera (0, 1000, pi/2, 0)
    assert d.metric == A.METRIC_KERR and d.metric_params[0] == 0.5 and d.rlog_scale == 0
    assert d.diag_flags == FCS and d.value_diag == FCS
    assert d.term_flags == A.TERMFLAG_BOUNDARYSPHERE | A.TERMFLAG_TIMEOUT
    assert d.sphere_radius == 1000.0 and d.timeout_max_steps == 10000
    assert list(d.camera_position) == [0.0, 1000.0, 3.1415926535 / 2.0, 0.0]
    assert list(d.screen_size) == [10.0, 10.0]
    m = _cfg("").mesh_desc()
    assert m.type == P.MESH_SIMPLESQUARE and m.total_pixels == 10000


def test_typed_lookups_follow_the_reference():
    text = """
    Metric = { Name = "kErR"; a = 1; RLogScale = true; };   // a given as an integer: ignored
    Terminations = {
      BoundarySphere = { On = true; SphereRadius = 500; };    /* integer: ignored, 1000 kept */
      TimeOut = { On = true; MaxSteps = 2000.0; };            # float for a counter: ignored
      Horizon = { On = true; Epsilon_Horizon = 2e-2; };
    };
    Integrator = { Type = "Verlet"; StepSize = 0.01; VerletVelocityTolerance = 1e-4; DerivativeH = 2e-7; };
    """
    d = _cfg(text).run_desc()
    assert d.metric == A.METRIC_KERR and d.metric_params[0] == 0.5 and d.rlog_scale == 1
    assert d.sphere_radius == 1000.0 and d.timeout_max_steps == 10000 and d.horizon_epsilon == 0.02
    assert d.term_flags == A.TERMFLAG_BOUNDARYSPHERE | A.TERMFLAG_TIMEOUT | A.TERMFLAG_HORIZON
    assert d.integrator == A.INTEGRATOR_VERLET and d.step_size == 0.01
    assert d.verlet_tolerance == 1e-4 and d.derivative_h == 2e-7


def test_syntax_variants():
    text = """
    # colon assignments, commas, adjacent strings, 64-bit integers, hex, lists
    Metric : { Name = "Rasheed" "-Larsen", m = 1.0, a = 0.6, p = 2.5, q = 3.0 }
    ViewScreen = {
      Mesh = { Type = "SquareSubdivisionMeshV2"; InitialPixels = 0x64; MaxPixels = 5000000000L;
               IterationPixels = 7; MaxSubdivide = 3; InitialSubdivisionToFinal = TRUE; };
      Extra = ( 1, 2.0, "three", { x = 1; } );
      Array = [ 1, 2, 3 ];
    };
    """
    c = _cfg(text)
    d, m = c.run_desc(), c.mesh_desc()
    assert d.metric == A.METRIC_RASHEEDLARSEN and list(d.metric_params[:4]) == [1.0, 0.6, 2.5, 3.0]
    assert m.type == P.MESH_SQUARESUBDIVISIONV2 and m.initial_pixels == 100
    assert m.max_pixels == 5000000000 and m.iteration_pixels == 7 and m.max_subdivide == 3
    assert m.initial_subdivision_to_final == 1


@pytest.mark.parametrize("bad", ['Metric = { Name = "Kerr" ', "a = ;", "a = 1; a = 2;", '"x" = 1;', "a = 3.14/2.0;"])
def test_syntax_errors_raise(bad):
    with pytest.raises(P.HostError):
        _cfg(bad)


def test_flat_space_turns_horizon_off_and_ignores_rlog():
    d = _cfg('Metric = { Name = "FlatSpace"; RLogScale = true; }; '
             'Terminations = { Horizon = { On = true; }; TimeOut = { On = true; }; };').run_desc()
    assert d.metric == A.METRIC_FLATSPACE and d.rlog_scale == 0
    assert d.term_flags == A.TERMFLAG_TIMEOUT


def test_counters_and_mesh_fallbacks():
    c = _cfg('ViewScreen = { Mesh = { Type = "SquareSubdivisionMesh"; InitialPixels = 400; MaxPixels = 10; '
             'MaxSubdivide = 0; IterationPixels = -5; }; }; '
             'Output = { FilePrefix = "x"; GeodesicsPerFile = 0; GeodesicsToCache = -1; };')
    m = c.mesh_desc()
    # MaxPixels below InitialPixels is raised to it; MaxSubdivide < 1 becomes 1; a negative
    # 32-bit integer reads as 0 through the unsigned getter
    assert m.max_pixels == 400 and m.max_subdivide == 1 and m.iteration_pixels == 0
    m = _cfg('ViewScreen = { Mesh = { Type = "simplesquaremesh"; }; };').mesh_desc()
    assert m.type == P.MESH_SIMPLESQUARE and m.total_pixels == 10000  # the type name is case-sensitive


def test_value_diagnostic_selection_and_order():
    both = 'Diagnostics = { FourColorScreen = { On = true; }; EquatorialPasses = { On = true; UseForMesh = true; }; };'
    c = _cfg(both)
    assert c.run_desc().value_diag == EP and c.diag_names() == ["EquatPasses", "FourColorScreen"]
    three = ('Diagnostics = { FourColorScreen = { On = true; }; EquatorialPasses = { On = true; }; '
             'ClosestRadius = { On = true; UseForMesh = true; }; };')
    # the value diagnostic is rotated with std::rotate(begin, begin + 1, end) as it is added, which
    # for three diagnostics puts EquatorialPasses first (FOORT/Diagnostics.cpp:14-99)
    c = _cfg(three)
    assert c.run_desc().value_diag == CR and c.diag_names() == ["EquatPasses", "ClosestRadius", "FourColorScreen"]
    none = 'Diagnostics = { EquatorialPasses = { On = true; }; };'
    d = _cfg(none).run_desc()  # no mesh diagnostic and no FourColorScreen: everything falls back
    assert d.diag_flags == FCS and d.value_diag == FCS


# ---------------------------------------------------------------------------
# 2. first line of the output files
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("run", ["test", "debug"])
def test_first_line_matches_the_shipped_output_files(run):
    g = np.load(os.path.join(GOLD, "reference_outputs.npz"))
    c = P.Config(path=os.path.join(GOLD, f"johannsen_{run}.cfg"))
    key = "Test" if run == "test" else "debug"
    assert c.first_line() == str(g[f"{key}_fcs_header"]) == str(g[f"{key}_ep_header"])
    assert c.diag_names() == ["EquatPasses", "FourColorScreen"]


FIRST_LINE_CASES = [
    dict(metric=A.METRIC_KERR, params=(0.94,), rlog=True, integrator=A.INTEGRATOR_RK4, diags=FCS | EP | CR,
         value_diag=CR, mesh=(P.MESH_SQUARESUBDIVISIONV2, 400, 9000, 50, 5, True)),
    dict(metric=A.METRIC_FLATSPACE, params=(), rlog=False, integrator=A.INTEGRATOR_VERLET, diags=CR,
         value_diag=CR, mesh=(P.MESH_SQUARESUBDIVISION, 100, 0, 10, 2, False)),
    dict(metric=A.METRIC_RASHEEDLARSEN, params=(1.0, 0.6, 2.5, 3.0), rlog=True, integrator=A.INTEGRATOR_RK4,
         diags=EP | CR, value_diag=EP, mesh=(P.MESH_SIMPLESQUARE, 1000, 0, 0, 1, False)),
    dict(metric=A.METRIC_MANKONOVIKOV, params=(0.3, 2.0), rlog=False, integrator=A.INTEGRATOR_VERLET,
         diags=FCS | EP, value_diag=FCS, mesh=(P.MESH_SIMPLESQUARE, 62500, 0, 0, 1, False)),
    dict(metric=A.METRIC_JOHANNSEN, params=(0.5, 1.0, 0.3, -0.2, 0.1), rlog=True, integrator=A.INTEGRATOR_RK4,
         diags=FCS | CR, value_diag=CR, mesh=(P.MESH_SQUARESUBDIVISION, 2500, 100000, 999, 7, False)),
]


def _cfg_text(case, terms_text, extra=""):
    names = {A.METRIC_KERR: "Kerr", A.METRIC_FLATSPACE: "FlatSpace", A.METRIC_RASHEEDLARSEN: "Rasheed-Larsen",
             A.METRIC_JOHANNSEN: "Johannsen", A.METRIC_MANKONOVIKOV: "Manko-Novikov"}
    keys = {A.METRIC_KERR: ["a"], A.METRIC_FLATSPACE: [], A.METRIC_RASHEEDLARSEN: ["m", "a", "p", "q"],
            A.METRIC_JOHANNSEN: ["a", "alpha13", "alpha22", "alpha52", "epsilon3"],
            A.METRIC_MANKONOVIKOV: ["a", "alpha3"]}
    ps = "".join(f" {k} = {float(v)!r};" for k, v in zip(keys[case["metric"]], case["params"]))
    diag = ""
    for flag, name in ((FCS, "FourColorScreen"), (EP, "EquatorialPasses"), (CR, "ClosestRadius")):
        if case["diags"] & flag:
            diag += f" {name} = {{ On = true; UseForMesh = {'true' if case['value_diag'] == flag else 'false'};"
            diag += " Threshold = 0.02; };" if flag == EP else " };"
    mt, pixels, maxp, it, ms, tf = case["mesh"]
    mname = {P.MESH_SIMPLESQUARE: "SimpleSquareMesh", P.MESH_SQUARESUBDIVISION: "SquareSubdivisionMesh",
             P.MESH_SQUARESUBDIVISIONV2: "SquareSubdivisionMeshV2"}[mt]
    mesh = (f'Type = "{mname}"; TotalPixels = {pixels};' if mt == P.MESH_SIMPLESQUARE else
            f'Type = "{mname}"; InitialPixels = {pixels}; MaxPixels = {maxp}; IterationPixels = {it}; '
            f'MaxSubdivide = {ms}; InitialSubdivisionToFinal = {"true" if tf else "false"};')
    integ = "RK4" if case["integrator"] == A.INTEGRATOR_RK4 else "Verlet"
    return f"""
    Metric = {{ Name = "{names[case['metric']]}";{ps} RLogScale = {'true' if case['rlog'] else 'false'}; }};
    Diagnostics = {{{diag} }};
    Terminations = {{ {terms_text} }};
    ViewScreen = {{ Position = {{ r = 500.0; theta = 1.2; phi = 0.25; }}; ScreenSize = {{ x = 20.0; y = 16.0; }};
                    ScreenCenter = {{ x = 1.0; y = -0.5; }}; Mesh = {{ {mesh} }}; }};
    Integrator = {{ Type = "{integ}"; StepSize = 0.02; VerletVelocityTolerance = 0.002; }};
    {extra}
    """


TERMS_TEXT = ("Horizon = { On = true; Epsilon_Horizon = 0.05; }; BoundarySphere = { On = true; SphereRadius = 900.0; };"
              " TimeOut = { On = true; MaxSteps = 50000; }; ThetaSingularity = { On = true; Epsilon = 1e-6; };")


@need_ref
@pytest.mark.parametrize("i", range(len(FIRST_LINE_CASES)))
def test_first_line_matches_reference(i):
    case = FIRST_LINE_CASES[i]
    c = _cfg(_cfg_text(case, TERMS_TEXT))
    d, m = c.run_desc(), c.mesh_desc()
    # the mesh as the configuration built it (MaxPixels below InitialPixels, 0 included, is raised to it)
    pixels = m.total_pixels if m.type == P.MESH_SIMPLESQUARE else m.initial_pixels
    ref = O.ref_first_line(d, m.type, pixels, m.max_pixels, m.iteration_pixels, m.max_subdivide,
                           m.initial_subdivision_to_final)
    assert c.first_line() == ref


# ---------------------------------------------------------------------------
# 3. meshes against the reference's, fed identical values
# ---------------------------------------------------------------------------

def _value_grid(kind, side, seed):
    rng = np.random.default_rng(seed)
    r, c = np.meshgrid(np.arange(side), np.arange(side), indexing="ij")
    x = (r - side * rng.uniform(0.3, 0.7)) / side * 10
    y = (c - side * rng.uniform(0.3, 0.7)) / side * 10
    rad = np.hypot(x, y)
    if kind == "fcs":  # a shadow disc on a four-colour sky: many equal weights
        return np.where(rad < 2.6, 0, np.where(x > 0, 1, 2) + np.where(y > 0, 0, 2)).astype(float)
    if kind == "ep":  # rings of equatorial passes, negative inside
        return np.where(rad < 2.0, -np.floor(rad * 2), np.floor(3 / (rad - 1.5))).astype(float)
    return np.round(rad * 3 + np.sin(x * 3) * 0.5, 3)  # continuous closest radii


def _drive(mesh, grid):
    seq, its = [], []
    while not mesh.is_finished:
        si = mesh.screen_indices()
        its.append(len(si))
        seq.append(si)
        mesh.finished(grid[si[:, 0].astype(np.int64), si[:, 1].astype(np.int64)])
        mesh.end_loop()
    return (np.concatenate(seq) if seq else np.zeros((0, 2), np.uint64)), its


MESH_CASES = [
    # (type, initial, max pixels, iteration pixels, max subdivide, to final, value kind, value diagnostic)
    (P.MESH_SQUARESUBDIVISION, 100, 2000, 40, 4, False, "fcs", FCS),
    (P.MESH_SQUARESUBDIVISIONV2, 100, 2000, 40, 4, False, "fcs", FCS),
    (P.MESH_SQUARESUBDIVISION, 256, 3000, 25, 5, True, "ep", EP),
    (P.MESH_SQUARESUBDIVISIONV2, 256, 3000, 25, 5, True, "ep", EP),
    (P.MESH_SQUARESUBDIVISION, 64, 0, 30, 3, False, "cr", CR),      # no pixel limit: runs to MaxSubdivide
    (P.MESH_SQUARESUBDIVISIONV2, 64, 0, 30, 3, False, "cr", CR),
    (P.MESH_SQUARESUBDIVISIONV2, 1000, 6000, 300, 4, False, "fcs", FCS),  # InitialPixels not a square
    (P.MESH_SQUARESUBDIVISION, 1000, 6000, 300, 4, False, "fcs", FCS),
    (P.MESH_SQUARESUBDIVISIONV2, 400, 400, 10, 3, False, "fcs", FCS),   # MaxPixels == InitialPixels
    (P.MESH_SQUARESUBDIVISIONV2, 49, 100000, 1000, 6, True, "cr", CR),
    (P.MESH_SQUARESUBDIVISION, 49, 100000, 1000, 6, True, "cr", CR),
    (P.MESH_SQUARESUBDIVISIONV2, 100, 50000, 5, 9, False, "fcs", FCS),  # few subdivisions per iteration
    (P.MESH_SIMPLESQUARE, 1000, 0, 0, 1, False, "fcs", FCS),
]


@need_ref
@pytest.mark.parametrize("case", range(len(MESH_CASES)))
@pytest.mark.parametrize("seed", [0, 1])
def test_mesh_queues_match_reference(case, seed):
    mt, init, maxp, it, ms, tf, kind, vd = MESH_CASES[case]
    md = P.mesh_desc(mt, vd, total_pixels=init, initial_pixels=init, max_pixels=maxp, iteration_pixels=it,
                     max_subdivide=ms, initial_subdivision_to_final=tf)
    mesh = P.Mesh(md)
    grid = _value_grid(kind, int(mesh.side), seed)
    ref_seq, ref_its, ref_desc = O.ref_mesh_run(mt, init, maxp, it, ms, tf, vd, grid)
    seq, its = _drive(P.Mesh(md), grid)
    assert mesh.description() == ref_desc
    assert its == [int(v) for v in ref_its]
    assert seq.shape == ref_seq.shape and (seq == ref_seq).all()
    if maxp:
        assert len(seq) <= max(maxp, (int(np.sqrt(init))) ** 2)


def test_mesh_rejects_bad_descriptions():
    with pytest.raises(P.HostError):
        P.Mesh(P.mesh_desc(P.MESH_SQUARESUBDIVISION, FCS, initial_pixels=100, max_subdivide=0))
    with pytest.raises(P.HostError):
        P.Mesh(P.mesh_desc(99, FCS, total_pixels=100))


def test_input_certain_pixels_mesh():
    px = np.array([[3, 4], [0, 0], [9, 9]], np.uint64)
    md = P.mesh_desc(P.MESH_INPUTCERTAINPIXELS, FCS, total_pixels=100)
    md.pixels = px.ctypes.data_as(C.c_void_p)
    md.pixel_count = 3
    m = P.Mesh(md)
    assert m.count == 3 and m.side == 10 and (m.screen_indices() == px).all()
    m.finished(np.zeros(3))
    m.end_loop()
    assert m.is_finished and m.description() == "Mesh: User-input pixels"


INPUT_PIXELS_CFG = """
Metric = {{ Name = "Kerr"; a = 0.6; RLogScale = true; }};
Diagnostics = {{ FourColorScreen = {{ On = true; UseForMesh = true; }}; EquatorialPasses = {{ On = true; }};
                ClosestRadius = {{ On = true; }}; }};
Terminations = {{ Horizon = {{ On = true; }}; BoundarySphere = {{ On = true; SphereRadius = 1000.0; }};
                 TimeOut = {{ On = true; MaxSteps = 20000; }}; }};
ViewScreen = {{ Position = {{ r = 1000.0; theta = 1.2; }}; ScreenSize = {{ x = 20.0; y = 20.0; }};
               Mesh = {{ Type = "InputCertainPixelsMesh"; TotalPixels = 100; }}; }};
Integrator = {{ Type = "RK4"; }};
Output = {{ FilePrefix = "{prefix}"; FileExtension = "dat"; TimeStamp = false; ScreenOutputLevel = 1; }};
"""
# pixels (3, 4) and (0, 0), then (99, 1) outside the 10 x 10 screen, then a negative number to stop
INPUT_PIXELS_STDIN = "3 4\n0 0\n99 1\n-1\n"

_RUN_ORACLE_PROGRAM = """
import sys
sys.path.insert(0, {repo!r})
import oracle as O
from foort_amd import program as P
P.run({cfg!r}, tracer=lambda d, s, si: O.trace("fm", d, s, screen_index=si))
"""
_RUN_REFERENCE_PROGRAM = """
import sys
sys.path.insert(0, {repo!r})
import oracle as O
from foort_amd import program as P
c = P.Config(path={cfg!r})
O.ref_program_run(c.run_desc(), P.MESH_INPUTCERTAINPIXELS, c.mesh_desc().total_pixels, {prefix!r}, ext="dat")
"""


def _dialogue(text):
    return [ln for ln in text.splitlines() if "InputCertainPixelsMesh message" in ln or "No pixels added" in ln]


@need_ref
def test_input_certain_pixels_dialogue_matches_reference(tmp_path):
    """InputCertainPixelsMesh's console dialogue (FOORT/Mesh.cpp:81-141) and the files of the
    run: the program (oracle as its batch tracer) against the reference's own mesh class and
    program loop (oracle/_ref), the same console input piped into both."""
    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    cfg = tmp_path / "pixels.cfg"
    cfg.write_text(INPUT_PIXELS_CFG.format(prefix=str(tmp_path / "new" / "run")))
    (tmp_path / "new").mkdir()
    (tmp_path / "ref").mkdir()
    new = subprocess.run([sys.executable, "-c", _RUN_ORACLE_PROGRAM.format(repo=repo, cfg=str(cfg))],
                         input=INPUT_PIXELS_STDIN, capture_output=True, text=True, timeout=120)
    assert new.returncode == 0, new.stderr
    ref = subprocess.run([sys.executable, "-c", _RUN_REFERENCE_PROGRAM.format(repo=repo, cfg=str(cfg),
                                                                                prefix=str(tmp_path / "ref" / "run"))],
                         input=INPUT_PIXELS_STDIN, capture_output=True, text=True, timeout=120)
    assert ref.returncode == 0, ref.stderr
    d_new, d_ref = _dialogue(new.stdout), _dialogue(ref.stdout)
    assert d_new == d_ref
    assert sum("added." in ln for ln in d_new) == 2 and sum("Invalid coordinates" in ln for ln in d_new) == 1
    assert _tree(tmp_path / "new") == _tree(tmp_path / "ref")
    lines = (tmp_path / "new" / "run_FourColorScreen.dat").read_text().splitlines()
    assert [ln.split()[:2] for ln in lines[1:]] == [["3", "4"], ["0", "0"]]


def test_input_certain_pixels_dialogue_ends_at_end_of_input(tmp_path):
    """End of input ends the dialogue (the reference, reading a failed stream as 0, would add
    (0, 0) for ever); with no pixel the run integrates nothing and says so."""
    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    cfg = tmp_path / "pixels.cfg"
    cfg.write_text(INPUT_PIXELS_CFG.format(prefix=str(tmp_path / "run")))
    for stdin, added in (("5 5\n", 1), ("", 0)):
        r = subprocess.run([sys.executable, "-c", _RUN_ORACLE_PROGRAM.format(repo=repo, cfg=str(cfg))],
                           input=stdin, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        d = _dialogue(r.stdout)
        assert sum("added." in ln for ln in d) == added
        assert ("No pixels added to integration list!" in r.stdout) == (added == 0)


# ---------------------------------------------------------------------------
# 4. output handler against the reference's
# ---------------------------------------------------------------------------

def _tree(root):
    out = {}
    for dp, _, fs in os.walk(root):
        for f in fs:
            p = os.path.join(dp, f)
            out[os.path.relpath(p, root)] = open(p, "rb").read()
    return out


OUTPUT_CASES = [
    # (batches, to_cache, per_file, first line, timestamp, extension, names)
    ([7, 5, 9], 2 ** 64 - 2, 2 ** 64 - 1, "Header line; with text", "", "dat", ["A", "B"]),
    ([7, 5, 9], 10, 4, "", "", "txt", ["A", "B", "C"]),
    ([12, 3, 8, 1], 0, 5, "first", "221203-101010", "", ["X"]),
    ([4, 4, 4], 8, 4, "h", "", "dat", ["EquatPasses", "FourColorScreen"]),
    ([20], 3, 7, "", "", "out", ["Q", "R"]),
    ([1, 1, 1, 1, 1], 2, 2, "x", "", "dat", ["A"]),
]


@need_ref
@pytest.mark.parametrize("i", range(len(OUTPUT_CASES)))
def test_output_files_match_reference(i, tmp_path):
    batches, cache, per_file, first, ts, ext, names = OUTPUT_CASES[i]
    ref_dir, my_dir = tmp_path / "ref", tmp_path / "mine"
    O.ref_output_run(str(ref_dir / "sub" / "run"), names, ts, ext, cache, per_file, first, batches)
    o = P.Output(str(my_dir / "sub" / "run"), names, timestamp=ts, extension=ext, to_cache=cache,
                 per_file=per_file, first_line=first)
    for b, n in enumerate(batches):
        o.prepare(n)
        for j in range(n):
            o.geodesic(j, b, j, [f"v{b}_{j}_{k}" for k in range(len(names))])
    o.finished()
    ref, mine = _tree(ref_dir), _tree(my_dir)
    assert sorted(ref) == sorted(mine)
    for k in ref:
        assert ref[k] == mine[k], k


def test_output_to_console(capfd):
    o = P.Output("", ["A", "B"])
    o.prepare(2)
    o.geodesic(0, 0, 0, ["1", "2"])
    o.geodesic(1, 0, 1, ["3", "4"])
    o.finished()
    assert capfd.readouterr().out.splitlines()[-6:] == ["0 0 ", "1", "2", "0 1 ", "3", "4"]


# ---------------------------------------------------------------------------
# 5. the whole program: driver + CPU oracle vs the reference program loop
# ---------------------------------------------------------------------------

def _oracle_tracer(desc, side, si):
    return O.trace("libm", desc, side, screen_index=si)


PROGRAM_CASES = [
    # (metric text, mesh, diagnostics (flags, value diag), integrator text)
    ('Name = "Kerr"; a = 0.9; RLogScale = true;', (P.MESH_SQUARESUBDIVISIONV2, 144, 700, 30, 3, False),
     (FCS | EP | CR, CR), 'Type = "RK4";'),
    ('Name = "Kerr"; a = 0.6; RLogScale = true;', (P.MESH_SQUARESUBDIVISION, 100, 500, 25, 4, True),
     (FCS | EP, FCS), 'Type = "Verlet"; VerletVelocityTolerance = 0.001;'),
    ('Name = "Johannsen"; a = 0.7; alpha13 = 2.0; RLogScale = true;', (P.MESH_SIMPLESQUARE, 400, 0, 0, 1, False),
     (EP | CR, EP), 'Type = "Verlet"; VerletVelocityTolerance = -1.0;'),
]


@need_ref
@pytest.mark.parametrize("i", range(len(PROGRAM_CASES)))
def test_program_matches_reference_program(i, tmp_path):
    metric, (mt, pixels, maxp, it, ms, tf), (diags, vd), integ = PROGRAM_CASES[i]
    dtext = ""
    for flag, name in ((FCS, "FourColorScreen"), (EP, "EquatorialPasses"), (CR, "ClosestRadius")):
        if diags & flag:
            dtext += f" {name} = {{ On = true; UseForMesh = {'true' if vd == flag else 'false'}; }};"
    mname = {P.MESH_SIMPLESQUARE: "SimpleSquareMesh", P.MESH_SQUARESUBDIVISION: "SquareSubdivisionMesh",
             P.MESH_SQUARESUBDIVISIONV2: "SquareSubdivisionMeshV2"}[mt]
    mesh = (f'Type = "{mname}"; TotalPixels = {pixels};' if mt == P.MESH_SIMPLESQUARE else
            f'Type = "{mname}"; InitialPixels = {pixels}; MaxPixels = {maxp}; IterationPixels = {it}; '
            f'MaxSubdivide = {ms}; InitialSubdivisionToFinal = {"true" if tf else "false"};')
    cfg = tmp_path / "run.cfg"
    cfg.write_text(f"""
    Metric = {{ {metric} }};
    Diagnostics = {{{dtext} }};
    Terminations = {{ Horizon = {{ On = true; }}; BoundarySphere = {{ On = true; SphereRadius = 1000.0; }};
                     TimeOut = {{ On = true; MaxSteps = 20000; }}; }};
    ViewScreen = {{ Position = {{ r = 1000.0; theta = 1.3; }}; ScreenSize = {{ x = 24.0; y = 24.0; }};
                    Mesh = {{ {mesh} }}; }};
    Integrator = {{ {integ} }};
    Output = {{ FilePrefix = "{tmp_path}/mine/run"; FileExtension = "dat"; TimeStamp = false;
               GeodesicsToCache = 150; GeodesicsPerFile = 400; ScreenOutputLevel = 0; }};
    """)
    P.run(str(cfg), tracer=_oracle_tracer)
    d = P.Config(path=str(cfg)).run_desc()
    O.ref_program_run(d, mt, pixels, str(tmp_path / "ref" / "run"), "dat", maxp, it, ms, tf, to_cache=150,
                      per_file=400)
    ref, mine = _tree(tmp_path / "ref"), _tree(tmp_path / "mine")
    assert sorted(ref) == sorted(mine) and len(ref) >= 2
    for k in ref:
        assert ref[k] == mine[k], k


def _oracle_tracer_gp(desc, side, si):
    if desc.diag_flags & A.DIAGFLAG_GEODESICPOSITION:
        return O.trace_positions("libm", desc, side, screen_index=si)
    return O.trace("libm", desc, side, screen_index=si)


GP_CASES = [
    # (GeodesicPosition settings, other diagnostics, mesh text, integrator)
    ("UpdateFrequency = 3; UpdateStart = false; UpdateFinish = true; OutputSteps = 7;",
     "FourColorScreen = { On = true; UseForMesh = true; };",
     'Type = "SimpleSquareMesh"; TotalPixels = 100;', 'Type = "RK4";'),
    ("UpdateFrequency = 40; OutputSteps = 0; UseForMesh = true;",
     "EquatorialPasses = { On = true; };",
     'Type = "SquareSubdivisionMeshV2"; InitialPixels = 64; MaxPixels = 300; IterationPixels = 20; '
     'MaxSubdivide = 3;', 'Type = "RK4";'),
    ("UpdateFrequency = 0; UpdateStart = true; UpdateFinish = true; OutputSteps = 1;",
     "ClosestRadius = { On = true; UseForMesh = true; }; FourColorScreen = { On = true; };",
     'Type = "SquareSubdivisionMesh"; InitialPixels = 49; MaxPixels = 200; IterationPixels = 15; '
     'MaxSubdivide = 3;', 'Type = "Verlet"; VerletVelocityTolerance = 0.001;'),
]


@need_ref
@pytest.mark.parametrize("i", range(len(GP_CASES)))
def test_program_with_geodesic_position_matches_reference_program(i, tmp_path):
    """GeodesicPosition (saved positions, decimation to OutputSteps, the "N ;; x x x x ..." text and,
    as a mesh value, (theta, phi) with the arc-length distance) against the reference program."""
    gp, other, mesh, integ = GP_CASES[i]
    cfg = tmp_path / "run.cfg"
    cfg.write_text(f"""
    Metric = {{ Name = "Kerr"; a = 0.8; RLogScale = true; }};
    Diagnostics = {{ GeodesicPosition = {{ On = true; {gp} }}; {other} }};
    Terminations = {{ Horizon = {{ On = true; }}; BoundarySphere = {{ On = true; SphereRadius = 1000.0; }};
                     TimeOut = {{ On = true; MaxSteps = 3000; }}; }};
    ViewScreen = {{ Position = {{ r = 1000.0; theta = 1.2; }}; ScreenSize = {{ x = 20.0; y = 20.0; }};
                    Mesh = {{ {mesh} }}; }};
    Integrator = {{ {integ} }};
    Output = {{ FilePrefix = "{tmp_path}/mine/run"; FileExtension = "dat"; TimeStamp = false;
               ScreenOutputLevel = 0; }};
    """)
    P.run(str(cfg), tracer=_oracle_tracer_gp)
    c = P.Config(path=str(cfg))
    d, m = c.run_desc(), c.mesh_desc()
    assert d.diag_flags & A.DIAGFLAG_GEODESICPOSITION
    O.ref_program_run(d, m.type, m.initial_pixels if m.type else m.total_pixels, str(tmp_path / "ref" / "run"), "dat",
                      m.max_pixels, m.iteration_pixels, m.max_subdivide, bool(m.initial_subdivision_to_final))
    ref, mine = _tree(tmp_path / "ref"), _tree(tmp_path / "mine")
    assert "run_GeodesicPosition.dat" in ref
    assert sorted(ref) == sorted(mine)
    for k in ref:
        assert ref[k] == mine[k], k


def test_geodesic_position_decimation_rule():
    """The oracle keeps what the reference's rule keeps: save i when i % (S // keep) == 0, the last
    save in place of the last kept one; checked against the rule applied to the full list of saves."""
    d = W.c0(params=(0.7,), rlog=True, inclination_deg=70.0, diags=A.DIAGFLAG_FOURCOLORSCREEN |
             A.DIAGFLAG_GEODESICPOSITION, max_steps=4000)
    si = np.array([[3, 4], [8, 8], [0, 15], [15, 2]], np.uint64)
    d.gp_update = A.UpdateFrequency(1, 0, 1)
    d.gp_output_steps = 0
    full = O.trace_positions("fm", d, 16, screen_index=si)
    assert np.array_equal(full["kept"], full["steps"] + 1)  # every call saves: Reset and each step
    for keep in (1, 2, 5, 33, 10 ** 6):
        d.gp_output_steps = keep
        got = O.trace_positions("fm", d, 16, screen_index=si)
        for p in range(len(si)):
            pts = full["points"][full["offsets"][p]:full["offsets"][p + 1]]
            S = len(pts)
            if keep < S:
                J = S // keep
                idx = list(range(0, S, J))
                if (S - 1) % J:
                    idx[-1] = S - 1
            else:
                idx = list(range(S))
            mine = got["points"][got["offsets"][p]:got["offsets"][p + 1]]
            assert np.array_equal(mine, pts[idx]), (keep, p)
    # update every 7th call, and only at the start and the end
    d.gp_output_steps = 0
    d.gp_update = A.UpdateFrequency(7, 1, 1)
    got = O.trace_positions("fm", d, 16, screen_index=si)
    for p in range(len(si)):
        S = int(full["kept"][p])
        assert np.array_equal(got["points"][got["offsets"][p]:got["offsets"][p + 1]],
                              full["points"][full["offsets"][p]:full["offsets"][p + 1]][6::7][:S // 7])
    d.gp_update = A.UpdateFrequency(0, 1, 1)
    got = O.trace_positions("fm", d, 16, screen_index=si)
    for p in range(len(si)):
        pts = full["points"][full["offsets"][p]:full["offsets"][p + 1]]
        assert np.array_equal(got["points"][got["offsets"][p]:got["offsets"][p + 1]], pts[[0, -1]])
    d.gp_update = A.UpdateFrequency(0, 0, 0)
    assert O.trace_positions("fm", d, 16, screen_index=si)["points"].shape == (0, 4)


# ---------------------------------------------------------------------------
# the command line program and the exported symbols
# ---------------------------------------------------------------------------

def test_cli_early_exits_like_the_reference(tmp_path):
    r = subprocess.run([P.CLI_PATH], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "No configuration file given" in r.stdout
    r = subprocess.run([P.CLI_PATH, str(tmp_path / "missing.cfg")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "Config file not found" in r.stderr
    bad = tmp_path / "bad.cfg"
    bad.write_text("Metric = { a = 3.14/2.0; };\n")
    r = subprocess.run([P.CLI_PATH, str(bad)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "parse error at" in r.stderr and ":1 -" in r.stderr


def test_host_library_exports_every_declared_symbol():
    decl = open(os.path.join(os.path.dirname(HERE), "include", "foort_host.h")).read()
    names = set(re.findall(r"\b(foort_[a-z_0-9]+)\s*\(", decl)) - {"foort_trace_fn", "foort_positions_fn"}
    assert names == set(P.EXPORTS)
    L = P.lib()
    for n in names:
        assert hasattr(L, n), n


def test_program_writes_value_texts_like_to_string(tmp_path):
    """The driver keeps the diagnostics as numbers and formats them when it writes the
    files; the texts must be std::to_string's: "%d" for the quadrant and the passes,
    "%f" (six decimals, correctly rounded) for the closest radius."""
    vals = np.array([0.0, -1.0, 2.5e-7, 5e-7, 1.5e-6, 0.1234565, 2.0000005, 999.9999995, 1e-300, 123456789.123456789,
                     1e22, 3.141592653589793, 0.3, 7.25, 1000.0, 6.0000004999], np.float64)
    side = 4

    def tracer(desc, s, si):
        n = len(si)
        assert s == side and n == side * side
        return {"term": np.full(n, 2, np.int32), "quadrant": np.arange(n, dtype=np.int32) % 5,
                "equat_passes": np.arange(n, dtype=np.int32) - 7, "closest_radius": vals[:n]}

    cfg = tmp_path / "v.cfg"
    cfg.write_text(f"""
    Diagnostics = {{ FourColorScreen = {{ On = true; UseForMesh = true; }}; EquatorialPasses = {{ On = true; }};
                    ClosestRadius = {{ On = true; }}; }};
    ViewScreen = {{ Mesh = {{ Type = "SimpleSquareMesh"; TotalPixels = {side * side}; }}; }};
    Output = {{ FilePrefix = "{tmp_path}/v"; FileExtension = "dat"; TimeStamp = false; FirstLineInfo = false;
               ScreenOutputLevel = 0; }};
    """)
    P.run(str(cfg), tracer=tracer)
    rows = [(i // side, i % side) for i in range(side * side)]
    cr = (tmp_path / "v_ClosestRadius.dat").read_text().splitlines()
    assert cr == [f"{r} {c}  {v:f}" for (r, c), v in zip(rows, vals)]
    ep = (tmp_path / "v_EquatPasses.dat").read_text().splitlines()
    assert ep == [f"{r} {c}  {i - 7}" for i, (r, c) in enumerate(rows)]


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present: the requests would run")
def test_server_without_gpu_returns_failed_requests():
    """Failure handling on a machine without a GPU: every request comes back with an error
    status (its group's device queue, or a GeodesicPosition request's own context, cannot be
    opened), drain and shutdown do not hang, and a request with a pixel list reaching off the
    screen is refused at submission."""
    reqs = W.c4_requests(seed=5, n=6, rate=1000.0)
    with P.Server(workers=2, hbm_budget_bytes=16 << 20) as srv:
        for k, (_, desc, mesh, _) in enumerate(reqs):
            srv.submit(desc, mesh, k)
        srv.drain()
        res = srv.results()
        assert sorted(r.tag for r in res) == list(range(6))
        assert all(r.status != 0 for r in res)
        d = reqs[0][1]
        d.diag_flags |= A.DIAGFLAG_GEODESICPOSITION
        srv.submit(d, reqs[0][2], 99)
        srv.drain()
        res = srv.results()
        assert [r.tag for r in res] == [99] and res[0].status != 0
        m = P.mesh_desc(P.MESH_INPUTCERTAINPIXELS, A.DIAGFLAG_FOURCOLORSCREEN, total_pixels=100,
                        pixels=[[0, 0], [9, 10]])
        with pytest.raises(P.HostError):
            srv.submit(reqs[1][1], m, 98)
