"""C2 at full size: the adaptive SquareSubdivisionMeshV2 run on Kerr a=0.94, 60 deg (256x256
initial grid, MaxSubdivide 9, at most 4,000,000 pixels; RLogScale on, so geodesic lengths
range from ~10^2 steps to the 1e5-step time-out), through the drop-in program.

The run with the persistent work-stealing schedule -- its options and the schedule picked by
the autotuner per batch size, the last geodesics of each batch finished by the four-lane tail
kernel, every mesh iteration one captured HIP graph -- writes the same output files, byte
for byte, as the plain one-shot run; and a sample of its pixels, taken from the files of
every mesh iteration, equals the CPU oracle's values for those pixels (the reference's
per-pixel loop, FOORT/Main.cpp:373-400).
"""
import hashlib
import os

import numpy as np
import pytest

import oracle as O
from foort_amd import program as P
from foort_amd import workloads as W

pytestmark = pytest.mark.gpu

NAMES = ("FourColorScreen", "EquatPasses", "ClosestRadius")


def _digest(path):
    h = hashlib.sha256()
    with open(path, "rb") as f:
        for chunk in iter(lambda: f.read(1 << 22), b""):
            h.update(chunk)
    return h.hexdigest()


def test_c2_full_size_tuned_persistent_graph_equals_oneshot_and_oracle(tmp_path, monkeypatch):
    monkeypatch.setenv("FOORT_TUNE_CACHE", str(tmp_path / "tune.txt"))
    runs = {}
    for name, sched, tune, graph in (("tuned", "Persistent", "Schedule", True), ("oneshot", "OneShot", "None", False)):
        cfg = tmp_path / f"{name}.cfg"
        cfg.write_text(W.c2_config(schedule=sched, autotune=tune, use_graph=graph, prefix=str(tmp_path / name)))
        print(f"[c2] {name}: schedule {sched}, autotune {tune}, graphs {graph}", flush=True)
        st = P.run(str(cfg), screen_output_level=1)
        print(f"[c2] {name}: {st.geodesics} pixels, {st.steps} steps, {st.launches} launches, "
              f"{st.device_ms / 1e3:.2f} s of kernels", flush=True)
        runs[name] = st
    assert runs["tuned"].geodesics == runs["oneshot"].geodesics > 2_000_000
    assert runs["tuned"].steps == runs["oneshot"].steps
    for n in NAMES:
        a, b = tmp_path / f"tuned_{n}.dat", tmp_path / f"oneshot_{n}.dat"
        assert _digest(a) == _digest(b), n

    # an oracle sample: every 80,000th line of the files (all mesh iterations)
    cfg = P.Config(path=str(tmp_path / "tuned.cfg"))
    desc, side = cfg.run_desc(), P.Mesh(cfg.mesh_desc()).side
    cols = []
    for n in NAMES:
        with open(tmp_path / f"tuned_{n}.dat") as f:
            f.readline()  # the run description
            cols.append([line.split() for k, line in enumerate(f) if k % 80000 == 0])
    si = np.array([[int(r), int(c)] for r, c, _ in cols[0]], np.uint64)
    print(f"[c2] oracle on {len(si)} pixels of the {side}^2 grid", flush=True)
    o = O.trace("fm", desc, side, screen_index=si)
    assert [int(v) for _, _, v in cols[0]] == o["quadrant"].tolist()
    assert [int(v) for _, _, v in cols[1]] == o["equat_passes"].tolist()
    # ClosestRadius is written as std::to_string of the double ("%f")
    assert [v for _, _, v in cols[2]] == [f"{x:f}" for x in o["closest_radius"]]
