"""The render server against the CPU oracle, on requests of the C4 stream (seed 2024).

A subset of the stream -- every metric kind, adaptive meshes of 2 to 6 iterations, final
sides up to 100 (what the oracle finishes in seconds) -- is submitted to the server at once,
each request writing the reference's output files (foort_request.output_prefix).  The same
request, written as a configuration file (program.config_text), then runs through the
drop-in program with the CPU oracle as its batch tracer (the reference's per-pixel loop,
FOORT/Main.cpp:313-419, mesh refinement FOORT/Mesh.cpp:1092-1187, output files
FOORT/InputOutput.cpp:129-332).

* finite differences, binary64: the server's files equal the oracle program's byte for byte
  (first line, every iteration's pixels in order, every value);
* mixed precision (the stream as the server benchmark runs it): on the pixels both runs
  integrated, the integer diagnostics agree but on a few pixels of the diagnostic
  boundaries, and the closest radii agree to the accuracy of the binary32 far field.
"""
import os

import numpy as np
import pytest

import oracle as O
from foort_amd import abi as A
from foort_amd import program as P
from foort_amd import workloads as W

pytestmark = pytest.mark.gpu

# stream indices: Kerr, Johannsen, Rasheed-Larsen and Manko-Novikov requests of final side <= 100
SUBSET = [7, 17, 30, 40, 51, 56, 66, 68, 83, 88, 109, 112, 23, 41, 198, 264, 285, 407, 95, 137, 331, 435, 52]
NAMES = ("FourColorScreen", "EquatPasses", "ClosestRadius")


def _requests():
    reqs = W.c4_requests(seed=2024, n=max(SUBSET) + 1)
    return [(k, reqs[k]) for k in SUBSET]


def _oracle_tracer(desc, side, si):
    return O.trace("fm", desc, side, screen_index=si)


def _oracle_files(tmp, k, desc, mesh):
    cfg = os.path.join(tmp, f"oracle_{k}.cfg")
    with open(cfg, "w") as f:
        f.write(P.config_text(desc, mesh, prefix=os.path.join(tmp, f"oracle_{k}")))
    P.run(cfg, screen_output_level=0, tracer=_oracle_tracer)
    return [os.path.join(tmp, f"oracle_{k}_{n}.dat") for n in NAMES]


def _serve(tmp, reqs, tag):
    with P.Server(workers=4, hbm_budget_bytes=64 << 20, launch_us=500) as srv:
        for k, (_, desc, mesh, _) in reqs:
            srv.submit(desc, mesh, k, output_prefix=os.path.join(tmp, f"{tag}_{k}"), output_extension="dat")
        srv.drain()
        res = {r.tag: r for r in srv.results(cap=len(reqs) + 4)}
    assert sorted(res) == sorted(k for k, _ in reqs)
    assert all(r.status == 0 for r in res.values())
    return res


def _values(path):
    """(row, column) -> value text of one output file (its first line is the run description)."""
    with open(path) as f:
        lines = f.read().splitlines()
    out = {}
    for line in lines[1:]:
        row, col, val = line.split()
        out[(int(row), int(col))] = val
    return lines[0], out


def test_server_fd_requests_equal_oracle_files(tmp_path):
    tmp = str(tmp_path)
    reqs = []
    for k, (t, desc, mesh, info) in _requests():
        desc.christoffel, desc.precision = A.CHRISTOFFEL_FD, A.PRECISION_FP64
        reqs.append((k, (t, desc, mesh, info)))
    kinds = {info["kind"] for _, (_, _, _, info) in reqs}
    assert kinds == {"kerr", "johannsen", "rasheed_larsen", "manko_novikov"}
    print(f"[server-oracle] {len(reqs)} finite-difference requests", flush=True)
    res = _serve(tmp, reqs, "server")
    for k, (_, desc, mesh, info) in reqs:
        want = _oracle_files(tmp, k, desc, mesh)
        for name, w in zip(NAMES, want):
            got = os.path.join(tmp, f"server_{k}_{name}.dat")
            with open(got, "rb") as a, open(w, "rb") as b:
                assert a.read() == b.read(), (k, info, name)
        lines = sum(1 for _ in open(want[0])) - 1
        assert res[k].pixels == lines, (k, info)
    print("[server-oracle] every file byte for byte", flush=True)


def test_server_mixed_requests_close_to_oracle(tmp_path):
    """An adaptive mesh refines where neighbouring values differ, so after the uniform initial
    grid nearly every pixel lies on a diagnostic boundary, where the formulations may differ:
    on the uniform initial grid of each request every mismatch lies on a diagnostic boundary
    (as in the C3 study), and over all common pixels of all requests at most 2% differ."""
    tmp = str(tmp_path)
    reqs = _requests()
    print(f"[server-oracle] {len(reqs)} mixed-precision requests", flush=True)
    _serve(tmp, reqs, "server")
    total = mismatched = 0
    for k, (_, desc, mesh, info) in reqs:
        fd = desc.__class__.from_buffer_copy(desc)
        fd.christoffel, fd.precision = A.CHRISTOFFEL_FD, A.PRECISION_FP64
        want = _oracle_files(tmp, k, fd, mesh)
        got = [_values(os.path.join(tmp, f"server_{k}_{n}.dat")) for n in NAMES]
        ref = [_values(w) for w in want]
        # the first initial_pixels lines of a file are the initial grid (files keep iteration order)
        initial = list(ref[0][1])[:mesh.initial_pixels]
        common = sorted(set(got[0][1]) & set(ref[0][1]))
        assert set(initial) <= set(got[0][1]), (k, info)
        # the adaptive meshes follow their own values: both runs share most later pixels
        assert len(common) >= 0.9 * len(ref[0][1]), (k, info)

        def mismatch(px):
            bad = np.zeros(len(px), bool)
            for g, r in zip(got[:2], ref[:2]):
                bad |= np.array([int(g[1][p]) != int(r[1][p]) for p in px])
            return bad

        mism, mism0 = mismatch(common), mismatch(initial)
        # the initial grid is uniform: a mismatch there must lie on a diagnostic boundary (a
        # 4-neighbour on the grid with other integer values in the oracle's run)
        rows = sorted({p[0] for p in initial})
        step = rows[1] - rows[0]
        ints = {p: (int(ref[0][1][p]), int(ref[1][1][p])) for p in initial}
        edge = np.array([any(q in ints and ints[q] != ints[p]
                             for q in ((p[0] - step, p[1]), (p[0] + step, p[1]), (p[0], p[1] - step),
                                       (p[0], p[1] + step))) for p in initial])
        rg = np.array([float(got[2][1][p]) for p in common])
        rr = np.array([float(ref[2][1][p]) for p in common])
        rel = np.abs(rg - rr)[~mism] / np.maximum(np.abs(rr[~mism]), 1.0)
        print(f"[server-oracle] request {k} ({info['kind']}, side {info['side']}, {info['iterations']} iterations): "
              f"initial grid {int(mism0.sum())} of {len(initial)} differ ({int((mism0 & ~edge).sum())} off the "
              f"boundary); all {len(common)} common pixels: "
              f"integer mismatch {mism.mean():.4f}, closest radius median rel {np.median(rel):.1e}", flush=True)
        assert not (mism0 & ~edge).any(), (k, info)
        # ClosestRadius is written with 6 decimals (std::to_string), about 1e-6 of r = 1..1000
        assert np.median(rel) <= 1e-4, (k, info)
        total += len(common)
        mismatched += int(mism.sum())
    print(f"[server-oracle] all requests: {mismatched} of {total} common pixels differ", flush=True)
    assert mismatched <= 0.02 * total


def test_server_input_pixel_requests_match_oracle(tmp_path):
    """InputCertainPixelsMesh requests (FOORT/Mesh.cpp:81-185: a pixel list, one iteration),
    finite differences in binary64: each file holds the list's pixels in its order, with the
    oracle's values (ClosestRadius printed as the reference prints it, std::to_string).  A list
    with a pixel off the screen is refused at submission."""
    from tests.cases import c0, sample_pixels
    tmp = str(tmp_path)
    side = 64
    reqs = []
    for k, (a, inc) in enumerate([(0.5, 80.0), (0.9, 30.0), (0.99, 85.0)]):
        desc = c0(params=(a,), inclination_deg=inc, rlog=True, max_steps=20000)
        px = sample_pixels(side, 40, seed=k)
        mesh = P.mesh_desc(P.MESH_INPUTCERTAINPIXELS, A.DIAGFLAG_FOURCOLORSCREEN, total_pixels=side * side,
                           pixels=px)
        reqs.append((k, desc, mesh, px))
    with P.Server(workers=2, hbm_budget_bytes=64 << 20, launch_us=500) as srv:
        for k, desc, mesh, px in reqs:
            srv.submit(desc, mesh, k, output_prefix=os.path.join(tmp, f"list_{k}"), output_extension="dat")
        bad = P.mesh_desc(P.MESH_INPUTCERTAINPIXELS, A.DIAGFLAG_FOURCOLORSCREEN, total_pixels=side * side,
                          pixels=[[3, 4], [side, 0]])
        with pytest.raises(P.HostError):
            srv.submit(reqs[0][1], bad, 99)
        srv.drain()
        res = {r.tag: r for r in srv.results(cap=8)}
    assert sorted(res) == [k for k, _, _, _ in reqs]
    for k, desc, mesh, px in reqs:
        assert res[k].status == 0 and res[k].pixels == len(px) and res[k].iterations == 1, k
        o = O.trace("fm", desc, side, screen_index=px)
        want = {"FourColorScreen": [str(int(v)) for v in o["quadrant"]],
                "EquatPasses": [str(int(v)) for v in o["equat_passes"]],
                "ClosestRadius": ["%f" % v for v in o["closest_radius"]]}
        for name in NAMES:
            with open(os.path.join(tmp, f"list_{k}_{name}.dat")) as f:
                rows = [line.split() for line in f.read().splitlines()[1:]]
            assert [(int(r), int(c)) for r, c, _ in rows] == [tuple(map(int, p)) for p in px], (k, name)
            assert [v for _, _, v in rows] == want[name], (k, name)
    print(f"[server-oracle] {len(reqs)} pixel-list requests: files equal the oracle's values", flush=True)


GP_TEXT = """
Metric = {{ Name = "Kerr"; a = 0.8; RLogScale = true; }};
Diagnostics = {{ GeodesicPosition = {{ On = true; UpdateFrequency = 20; OutputSteps = 12; UseForMesh = {gp_mesh}; }};
                EquatorialPasses = {{ On = true; UseForMesh = {ep_mesh}; }}; }};
Terminations = {{ Horizon = {{ On = true; }}; BoundarySphere = {{ On = true; SphereRadius = 1000.0; }};
                 TimeOut = {{ On = true; MaxSteps = 3000; }}; }};
ViewScreen = {{ Position = {{ r = 1000.0; theta = 1.2; }}; ScreenSize = {{ x = 20.0; y = 20.0; }};
               Mesh = {{ Type = "{mesh}"; InitialPixels = 100; MaxPixels = 500; IterationPixels = 30;
                        MaxSubdivide = 3; TotalPixels = 400; }}; }};
Integrator = {{ Type = "RK4"; }};
"""


def _positions_tracer(desc, side, si):
    if desc.diag_flags & A.DIAGFLAG_GEODESICPOSITION:
        return O.trace_positions("fm", desc, side, screen_index=si)
    return O.trace("fm", desc, side, screen_index=si)


def test_server_geodesic_position_requests_equal_oracle_files(tmp_path):
    """Requests with GeodesicPosition (FOORT/Diagnostics.cpp:239-335) -- as the mesh value on an
    adaptive mesh and beside EquatorialPasses on a grid -- run whole on a worker thread next to
    a queue request; their files equal those of the oracle-driven program byte for byte."""
    tmp = str(tmp_path)
    reqs = []
    for k, (mesh, gp_mesh) in enumerate((("SquareSubdivisionMeshV2", "true"), ("SimpleSquareMesh", "false"))):
        c = P.Config(text=GP_TEXT.format(mesh=mesh, gp_mesh=gp_mesh, ep_mesh="false" if gp_mesh == "true" else "true"))
        reqs.append((k, c.run_desc(), c.mesh_desc(), c.diag_names()))
    queued = _requests()[0]
    with P.Server(workers=3, hbm_budget_bytes=64 << 20, launch_us=500) as srv:
        for k, desc, mesh, _ in reqs:
            srv.submit(desc, mesh, k, output_prefix=os.path.join(tmp, f"gp_{k}"), output_extension="dat")
        srv.submit(queued[1][1], queued[1][2], 7)
        srv.drain()
        res = {r.tag: r for r in srv.results(cap=8)}
    assert sorted(res) == [0, 1, 7] and all(r.status == 0 for r in res.values())
    for k, desc, mesh, names in reqs:
        assert "GeodesicPosition" in names
        cfg = os.path.join(tmp, f"oracle_gp_{k}.cfg")
        with open(cfg, "w") as f:
            f.write(P.config_text(desc, mesh, prefix=os.path.join(tmp, f"oracle_gp_{k}")))
        P.run(cfg, screen_output_level=0, tracer=_positions_tracer)
        for name in names:
            with open(os.path.join(tmp, f"gp_{k}_{name}.dat"), "rb") as a, \
                    open(os.path.join(tmp, f"oracle_gp_{k}_{name}.dat"), "rb") as b:
                assert a.read() == b.read(), (k, name)
        assert res[k].iterations >= 1 and res[k].pixels == sum(1 for _ in open(os.path.join(tmp, f"gp_{k}_{names[0]}.dat"))) - 1
    print(f"[server-oracle] {len(reqs)} GeodesicPosition requests: files byte for byte", flush=True)
