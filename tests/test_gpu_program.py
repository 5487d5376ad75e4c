"""The drop-in program on the GPU (`foort <file.cfg>`, libfoort_host.so over libfoort_gpu.so).

1. The command line program on the settings of the reference's two shipped
   runs writes FOORT/Output/{Test,debug}_FourColorScreen.dat and
   _EquatPasses.dat byte for byte (expected bytes rebuilt from
   tests/golden/reference_outputs.npz: the recorded first line, then
   "row col  value" for the 250 x 250 grid).
2. With an adaptive mesh the program writes the same files on the GPU as with
   the CPU oracle (portable-math build) as its batch tracer: the GPU path is bit
   for bit the oracle, so the mesh sees the same values and queues the same
   pixels in every iteration.
3. The persistent schedule gives the same files as the one-shot schedule.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
from foort_amd import program as P

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def _tree(root):
    out = {}
    for dp, _, fs in os.walk(root):
        for f in fs:
            p = os.path.join(dp, f)
            out[os.path.relpath(p, root)] = open(p, "rb").read()
    return out


def _expected(header, arr):
    side = arr.shape[0]
    lines = [header] + [f"{r} {c}  {int(arr[r, c])}" for r in range(side) for c in range(side)]
    return ("\n".join(lines) + "\n").encode()


@pytest.mark.parametrize("run", ["Test", "debug"])
def test_cli_reproduces_the_shipped_output_files(run, tmp_path):
    print(f"[program] foort johannsen_{run.lower()}.cfg (250x250 Johannsen)", flush=True)
    cfg = os.path.join(GOLD, f"johannsen_{run.lower()}.cfg")
    r = subprocess.run([P.CLI_PATH, cfg, "--output-prefix", str(tmp_path / run), "--screen-output-level", "1"],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout[-800:], flush=True)
    assert r.returncode == 0, r.stderr
    g = np.load(os.path.join(GOLD, "reference_outputs.npz"))
    assert sorted(os.listdir(tmp_path)) == [f"{run}_EquatPasses.dat", f"{run}_FourColorScreen.dat"]
    for diag, key in (("FourColorScreen", "fcs"), ("EquatPasses", "ep")):
        got = (tmp_path / f"{run}_{diag}.dat").read_bytes()
        want = _expected(str(g[f"{run}_{key}_header"]), g[f"{run}_{key}"])
        assert got == want, f"{diag}: {sum(a != b for a, b in zip(got.splitlines(), want.splitlines()))} lines differ"


ADAPTIVE = """
Metric = {{ Name = "Kerr"; a = 0.94; RLogScale = true; }};
Diagnostics = {{ FourColorScreen = {{ On = true; UseForMesh = true; }}; EquatorialPasses = {{ On = true; }};
                ClosestRadius = {{ On = true; }}; }};
Terminations = {{ Horizon = {{ On = true; }}; BoundarySphere = {{ On = true; SphereRadius = 1000.0; }};
                 TimeOut = {{ On = true; MaxSteps = 100000; }}; }};
ViewScreen = {{ Position = {{ r = 1000.0; theta = 1.0471975511965976; }}; ScreenSize = {{ x = 24.0; y = 24.0; }};
               Mesh = {{ Type = "{mesh}"; InitialPixels = 1024; MaxPixels = 6000; IterationPixels = 150;
                        MaxSubdivide = 4; }}; }};
Integrator = {{ Type = "RK4"; }};
Output = {{ FilePrefix = "unused"; FileExtension = "dat"; TimeStamp = false; ScreenOutputLevel = 0; }};
"""


def _fm_tracer(desc, side, si):
    return O.trace("fm", desc, side, screen_index=si)


@pytest.mark.parametrize("mesh", ["SquareSubdivisionMesh", "SquareSubdivisionMeshV2"])
def test_adaptive_program_matches_oracle(mesh, tmp_path):
    print(f"[program] adaptive {mesh} GPU vs oracle", flush=True)
    cfg = tmp_path / "adaptive.cfg"
    cfg.write_text(ADAPTIVE.format(mesh=mesh))
    st = P.run(str(cfg), output_prefix=str(tmp_path / "gpu" / "run"))
    assert st.geodesics > 1024
    P.run(str(cfg), output_prefix=str(tmp_path / "cpu" / "run"), tracer=_fm_tracer)
    gpu, cpu = _tree(tmp_path / "gpu"), _tree(tmp_path / "cpu")
    assert sorted(gpu) == sorted(cpu) == ["run_ClosestRadius.dat", "run_EquatPasses.dat", "run_FourColorScreen.dat"]
    for k in gpu:
        assert gpu[k] == cpu[k], k
    n = len(gpu["run_FourColorScreen.dat"].splitlines()) - 1
    assert n == st.geodesics


def test_persistent_schedule_writes_the_same_files(tmp_path):
    print("[program] persistent vs one-shot schedule", flush=True)
    cfg = tmp_path / "adaptive.cfg"
    cfg.write_text(ADAPTIVE.format(mesh="SquareSubdivisionMeshV2"))
    P.run(str(cfg), output_prefix=str(tmp_path / "one" / "run"))
    P.run(str(cfg), output_prefix=str(tmp_path / "per" / "run"), schedule=1)
    one, per = _tree(tmp_path / "one"), _tree(tmp_path / "per")
    assert one == per


def test_server_requests_match_direct_runs():
    """The render server runs requests concurrently: those of one metric share a device queue
    (time-sliced launches of 300 us, so geodesics span launches); every request's pixel and
    step counts equal those of the same adaptive mesh driven directly, one batch at a time,
    through a fresh context."""
    import foort_amd as F
    from foort_amd import workloads as W
    reqs = W.c4_requests(seed=7, n=12, rate=1000.0)
    reqs = [r for r in reqs if r[3]["side"] <= 300][:8]
    print(f"[server] {len(reqs)} requests", flush=True)
    with P.Server(workers=4, hbm_budget_bytes=64 << 20, launch_us=300) as srv:
        for k, (_, desc, mesh, _) in enumerate(reqs):
            srv.submit(desc, mesh, k)
        srv.drain()
        got = {r.tag: r for r in srv.results()}
    assert sorted(got) == list(range(len(reqs)))
    for k, (_, desc, mesh, info) in enumerate(reqs):
        r = got[k]
        assert r.status == 0 and 2 <= r.iterations <= info["iterations"]
        m = P.Mesh(mesh)
        pixels = steps = 0
        with F.Context(desc) as ctx:
            while not m.is_finished:
                si = m.screen_indices()
                out = ctx.trace_pixels(m.side, si)
                m.finished(out["quadrant"].astype(np.float64))
                m.end_loop()
                pixels += len(si)
                steps += int(out["stats"].steps)
        assert (r.pixels, r.steps) == (pixels, steps), (k, info)
        assert r.end_s >= r.start_s >= r.submit_s
        assert r.kernel_ms > 0


TUNED = """
Metric = {{ Name = "Kerr"; a = 0.9; RLogScale = true; }};
Diagnostics = {{ FourColorScreen = {{ On = true; UseForMesh = true; }}; EquatorialPasses = {{ On = true; }}; }};
Terminations = {{ Horizon = {{ On = true; }}; BoundarySphere = {{ On = true; SphereRadius = 1000.0; }};
                 TimeOut = {{ On = true; MaxSteps = 100000; }}; }};
ViewScreen = {{ Position = {{ r = 1000.0; theta = 1.2; }}; ScreenSize = {{ x = 24.0; y = 24.0; }};
               Mesh = {{ Type = "SimpleSquareMesh"; TotalPixels = 90000; }}; }};
Integrator = {{ Type = "RK4"; }};
Output = {{ FilePrefix = "unused"; FileExtension = "dat"; TimeStamp = false; ScreenOutputLevel = 0; }};
GPU = {{ Autotune = "{tune}"; UseGraph = true; }};
"""


def test_autotuned_program_writes_the_same_files(tmp_path, monkeypatch):
    """Tuning by batch size (batches of at least 65536 pixels; here one 300 x 300 grid) changes only
    how the batches run: the tuned run, the run that reuses the cached choice and an untuned run
    write the same files; the cache holds one entry per batch size tuned."""
    print("[program] autotuned run", flush=True)
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("FOORT_TUNE_CACHE", str(cache))
    runs = {}
    for name, tune in (("tuned", "Schedule"), ("cached", "Schedule"), ("plain", "None")):
        cfg = tmp_path / f"{name}.cfg"
        cfg.write_text(TUNED.format(tune=tune))
        st = P.run(str(cfg), output_prefix=str(tmp_path / name / "run"))
        runs[name] = _tree(tmp_path / name)
        print(f"[program] {name}: {st.geodesics} pixels, {st.device_ms:.0f} ms", flush=True)
    assert runs["tuned"] == runs["cached"] == runs["plain"]
    entries = cache.read_text().splitlines()
    assert entries and all("size=2^16" in e for e in entries)
    assert len(entries) == len({e.split(" | ")[0] for e in entries})  # the second run only read the cache


ERGO_CFG = """
Metric = { Name = "Kerr"; a = 0.99; RLogScale = false; };
Source = { Name = "NoSource"; };
Diagnostics = { FourColorScreen = { On = true; }; };
Terminations = {
  Horizon = { On = true; Epsilon_Horizon = 0.01; };
  BoundarySphere = { On = true; SphereRadius = 1000.0; };
  TimeOut = { On = true; MaxSteps = 2000; };
};
ViewScreen = {
  Position = { t = 0.0; r = 1.5; theta = 1.5707963267948966; phi = 0.0; };
  ScreenSize = { x = 4.0; y = 4.0; };
  Mesh = { Type = "SimpleSquareMesh"; TotalPixels = 64; };
};
Integrator = { Type = "RK4"; };
Output = { FilePrefix = "%s"; FileExtension = "dat"; TimeStamp = false; ScreenOutputLevel = 1; };
"""


def test_cli_prints_the_initial_velocity_warnings(tmp_path):
    """The program prints the reference's warning for every pixel whose initial velocity points
    to the future or outward (FOORT/ViewScreen.cpp:112-116): a camera at r = 1.5 in the
    ergoregion of Kerr a = 0.99, 8 x 8 pixels."""
    from tests.test_gpu_parity import _initial_velocities
    from foort_amd.workloads import c0
    cfg = tmp_path / "ergo.cfg"
    cfg.write_text(ERGO_CFG % str(tmp_path / "ergo"))
    print("[program] ergoregion camera", flush=True)
    r = subprocess.run([P.CLI_PATH, str(cfg)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    v = _initial_velocities(c0(params=(0.99,), inclination_deg=90.0, r_cam=1.5, screen=(4.0, 4.0)), 8)
    lines = r.stdout.splitlines()
    fut = "WARNING: Initial velocity of geodesic is future-pointing (should be past-pointing)!"
    out = "WARNING: Initial velocity of geodesic is outward-pointing (should be inward-pointing)!"
    assert lines.count(fut) == int((v[:, 0] > 0).sum()) > 0
    assert lines.count(out) == int((v[:, 1] > 0).sum()) > 0


def test_cli_input_certain_pixels_from_the_console(tmp_path):
    """`foort <cfg>` with InputCertainPixelsMesh reads its pixels from the console
    (FOORT/Mesh.cpp:81-141): the GPU program prints the dialogue and writes the same files as
    the program with the oracle as its batch tracer (tests/test_host.py pins that one to the
    reference's own mesh class)."""
    import sys
    from tests.test_host import INPUT_PIXELS_CFG, INPUT_PIXELS_STDIN, _RUN_ORACLE_PROGRAM, _dialogue
    print("[program] InputCertainPixelsMesh through the console", flush=True)
    repo = os.path.dirname(HERE)
    for side in ("gpu", "cpu"):
        (tmp_path / side).mkdir()
        cfg = tmp_path / f"{side}.cfg"
        cfg.write_text(INPUT_PIXELS_CFG.format(prefix=str(tmp_path / side / "run")))
        cmd = ([P.CLI_PATH, str(cfg)] if side == "gpu" else
               [sys.executable, "-c", _RUN_ORACLE_PROGRAM.format(repo=repo, cfg=str(cfg))])
        r = subprocess.run(cmd, input=INPUT_PIXELS_STDIN, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        if side == "gpu":
            d = _dialogue(r.stdout)
            assert sum("added." in ln for ln in d) == 2 and sum("Invalid coordinates" in ln for ln in d) == 1
    gpu, cpu = _tree(tmp_path / "gpu"), _tree(tmp_path / "cpu")
    assert sorted(gpu) == ["run_ClosestRadius.dat", "run_EquatPasses.dat", "run_FourColorScreen.dat"]
    for k in gpu:
        assert gpu[k] == cpu[k], k


PROGRESS = """
Metric = {{ Name = "Kerr"; a = 0.9; RLogScale = true; }};
Diagnostics = {{ FourColorScreen = {{ On = true; UseForMesh = true; }}; EquatorialPasses = {{ On = true; }}; }};
Terminations = {{ Horizon = {{ On = true; }}; BoundarySphere = {{ On = true; }}; TimeOut = {{ On = true; MaxSteps = 20000; }}; }};
ViewScreen = {{ Position = {{ r = 1000.0; theta = 1.3; }}; ScreenSize = {{ x = 24.0; y = 24.0; }};
               Mesh = {{ Type = "{mesh}"; {pixels} }}; }};
Integrator = {{ Type = "RK4"; }};
Output = {{ FilePrefix = "unused"; FileExtension = "dat"; TimeStamp = false; ScreenOutputLevel = 2; {freq} }};
"""


@pytest.mark.parametrize("mesh,pixels", [("SimpleSquareMesh", "TotalPixels = 262144;"),
                                         ("SquareSubdivisionMeshV2", "InitialPixels = 160000; MaxPixels = 400000; "
                                                                     "IterationPixels = 100000; MaxSubdivide = 3;")])
def test_loop_message_frequency_prints_progress_and_keeps_results(mesh, pixels, tmp_path):
    """Output.LoopMessageFrequency: the program prints the reference's progress line
    (FOORT/Main.cpp:354-371) between chunks of about that many geodesics, and the files are
    those of the run without it."""
    print(f"[program] LoopMessageFrequency, {mesh}", flush=True)
    outs = {}
    for name, freq in (("plain", ""), ("progress", "LoopMessageFrequency = 70000;")):
        cfg = tmp_path / f"{name}.cfg"
        cfg.write_text(PROGRESS.format(mesh=mesh, pixels=pixels, freq=freq))
        r = subprocess.run([P.CLI_PATH, str(cfg), "--output-prefix", str(tmp_path / name / "run")],
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        outs[name] = r.stdout
    lines = [ln for ln in outs["progress"].splitlines() if ln.startswith("Approx. at geodesic ")]
    print("\n".join(lines[:4]), flush=True)
    assert len(lines) >= 2 and not any(ln.startswith("Approx.") for ln in outs["plain"].splitlines())
    assert all(re.match(r"Approx\. at geodesic \d+ \(\d+\.\d+s elapsed; speed: \d+ geod/s; "
                        r"est\. loop time remaining: \d+\.\d+s\)\.\.\.$", ln) for ln in lines)
    assert _tree(tmp_path / "plain") == _tree(tmp_path / "progress")


@pytest.mark.parametrize("schedule", ["OneShot", "Persistent"])
def test_adaptive_run_with_graphs_launches_every_iteration_as_a_graph(schedule, tmp_path, capfd):
    """UseGraph on an adaptive mesh, whose batch sizes never repeat: every mesh iteration is a
    launch of one executable HIP graph updated in place, and the files equal those of the run
    without graphs."""
    print(f"[program] graphs on an adaptive V2 mesh, {schedule}", flush=True)
    cfg = tmp_path / "adaptive.cfg"
    base = ADAPTIVE.format(mesh="SquareSubdivisionMeshV2")
    runs = {}
    for graph in ("false", "true"):
        cfg.write_text(base + f'GPU = {{ Schedule = "{schedule}"; UseGraph = {graph}; }};\n')
        capfd.readouterr()
        runs[graph] = P.run(str(cfg), output_prefix=str(tmp_path / graph / "run"), screen_output_level=1)
        out = capfd.readouterr().out
    batches = len(re.findall(r"^Integrating [1-9]\d* geodesics on the GPU", out, re.M))
    plain, graphs = runs["false"], runs["true"]
    print(f"[program] {batches} mesh iterations, {graphs.graph_launches} graph launches", flush=True)
    assert batches >= 4
    assert plain.graph_launches == 0 and graphs.graph_launches == batches
    assert _tree(tmp_path / "false") == _tree(tmp_path / "true")
    assert graphs.steps == plain.steps and graphs.geodesics == plain.geodesics
