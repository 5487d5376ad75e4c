#!/usr/bin/env python3
"""bench.py -- throughput of the per-pixel null-geodesic integration on one MI355X.

Workload (the drop-in configuration): Kerr a=0.99, observer r=1000M,
inclination 85 deg, 1024x1024 SimpleSquareMesh over a 30M x 30M screen, RK4
with the adaptive step, FP64, the reference's finite-difference Christoffel
symbols, horizon / boundary sphere / time-out (1e5) terminations and the
FourColorScreen, EquatorialPasses and ClosestRadius diagnostics.  One
benchmark step = one full 1024^2 image (all pixels in one batch through the C
ABI, foort_gpu_trace_device), results resident in HBM; the pixel indices are
generated on the device, so no input crosses PCIe.

Reported: pixels/s (value) and geodesic steps/s of that path, whose results
are bit-identical to the CPU oracle; the roofline of its kernel; the other
formulations of the same image (persistent schedule, analytic Christoffels in
FP64 and mixed precision) with their agreement with it; and the reference's
own CPU build (oracle/_ref, its -Ofast flags) timed on a sample of the same
image on this host's cores.  The same line carries the metric's latency half:
the render server under the C4 request stream (2000 requests from seed 2024,
Poisson arrivals at 40/s, mixed precision), its p50/p99 latency and whether the
stream is sustained (--no-server skips it).

python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import platform
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
if REPO not in sys.path:
    sys.path.insert(0, REPO)

# The render server runs one device queue per metric group (four in the C4 stream), each
# time-sliced on its own stream; HIP maps streams onto GPU_MAX_HW_QUEUES hardware queues (4
# by default), and once a process has made other streams (the C1 contexts before the server
# leg) two groups can land on one hardware queue, where their kernels run one after the other
# (measured: Manko-Novikov p50 7.8 s -> 12.5 s).  HIP reads the variable when it initialises,
# before the first HIP call of this process.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or 4) < 8:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"

METRIC = "pixels/s & geodesic steps/s, Kerr a=0.99 1024² (FP64 vs mixed); p50/p99 latency"
PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 vector peak


def hbm_bytes_per_pixel(christoffel, precision, schedule):
    """HBM bytes per pixel of a path's kernels, measured with rocprofv3 FETCH_SIZE and WRITE_SIZE
    (each calibrated on a plain device copy) on the 1024x1024 image: foort_amd/hbm_traffic.json,
    written by foort_amd/counters.py; None for a path not measured."""
    from foort_amd import counters
    t = counters.load()
    e = t and t["paths"].get(f"{christoffel}/{precision}/{schedule}")
    return e["bytes_per_pixel"] if e else None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def flops_per_step(metric="kerr"):
    """Algorithmic FLOPs of one RK4 step: the closed-form (analytic) formulation
    counted by foort_amd/csrc/opcount.cpp (foort_amd/flops_per_step.json)."""
    with open(os.path.join(REPO, "foort_amd", "flops_per_step.json")) as f:
        d = json.load(f)
    closed = [c for c in d["counts"] if c["metric"] == metric and c["formulation"] != "fd"]
    best = min(closed, key=lambda c: c["rk4_step_flops"])
    fd = [c for c in d["counts"] if c["metric"] == metric and c["formulation"] == "fd"][0]
    return best, fd


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def host_threads():
    """Threads for the CPU baseline: the process's CPU share -- OMP_NUM_THREADS where the
    environment sets it (a GPU box gives each GPU's processes a share of its cores and says
    so there), else every core this process may run on."""
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    share = int(env) if env.isdigit() and int(env) > 0 else affinity
    return min(share, affinity), affinity


def cpu_baseline(desc, side, stride):
    """The reference's own path (oracle/_ref, built from its sources with its
    -Ofast flags) on every `stride`-th row and column of the image."""
    import numpy as np
    import oracle as O
    rows = np.arange(0, side, stride, dtype=np.uint64)
    px = np.stack(np.meshgrid(rows, rows, indexing="ij"), -1).reshape(-1, 2)
    threads, affinity = host_threads()
    log(f"cpu baseline on {threads} threads (cores available to this process: {affinity}, "
        f"os.cpu_count: {os.cpu_count()}, {cpu_model()})")
    kind = "reference" if O.have("ref_ofast") else "port"
    if kind == "reference":
        steps, sec = O.time_reference("ref_ofast", desc, side, px, threads)
    else:
        t = time.perf_counter()
        o = O.trace("libm", desc, side, screen_index=px, nthreads=threads)
        sec = time.perf_counter() - t
        steps = int(o["steps"].sum())
    return {
        "value": len(px) / sec, "unit": "pixels/s", "cores": threads, "kind": kind,
        "sample": f"{len(px)} pixels (every {stride}th row and column of the 1024x1024 image), "
                  f"{steps} geodesic steps in {sec:.1f} s on {threads} threads of {cpu_model()}",
        "geodesic_steps_per_s": steps / sec,
        "host": {"cpu_model": cpu_model(), "cores_available": affinity, "os_cpu_count": os.cpu_count(),
                 "omp_num_threads": os.environ.get("OMP_NUM_THREADS")},
    }


def agreement(ref, out):
    import numpy as np
    same = (ref["term"] == out["term"]) & (ref["quadrant"] == out["quadrant"]) & \
           (ref["equat_passes"] == out["equat_passes"])
    rel = np.abs(out["final_pos"] - ref["final_pos"]) / np.maximum(np.abs(ref["final_pos"]), 1e-300)
    relmax = rel.max(axis=1)[same]
    return {
        "integer_diag_mismatch_rate": float(1 - same.mean()),
        "pos_rel_err_median": float(np.median(relmax)) if relmax.size else None,
        "pos_rel_err_p99": float(np.quantile(relmax, 0.99)) if relmax.size else None,
        "frac_pos_within_1e-9": float((relmax <= 1e-9).mean()) if relmax.size else None,
    }


def quiet(fn):
    """Run fn with the process's stdout (file descriptor 1, which the C++ program
    writes its messages to) sent to stderr, so stdout carries only the JSON line."""
    sys.stdout.flush()
    saved = os.dup(1)
    os.dup2(2, 1)
    try:
        return fn()
    finally:
        os.dup2(saved, 1)
        os.close(saved)


def run_c2(args):
    """C2: the whole adaptive run through the drop-in program (configuration, mesh
    refinement on the host, one GPU batch per mesh iteration, output files in a
    temporary directory).  One step = one complete run."""
    import shutil
    import tempfile
    from foort_amd import program as P
    from foort_amd import workloads as W
    # (name, mesh, Christoffel, precision, schedule, autotune, graph); the first is the reported path:
    # the reference's finite differences (bit-exact FP64), schedule and options picked by the autotuner
    # on the first iteration, every mesh iteration one captured HIP graph
    V2, V1 = "SquareSubdivisionMeshV2", "SquareSubdivisionMesh"
    variants = [("fd_tuned_graph", V2, "FD", "FP64", "Persistent", "Schedule", True),
                ("fd_oneshot", V2, "FD", "FP64", "OneShot", "None", False),
                ("fd_persistent", V2, "FD", "FP64", "Persistent", "None", False),
                ("fd_persistent_graph", V2, "FD", "FP64", "Persistent", "None", True),
                ("analytic_fp64_tuned_graph", V2, "Analytic", "FP64", "Persistent", "Schedule", True),
                ("mixed_tuned_graph", V2, "Analytic", "Mixed", "Persistent", "Schedule", True),
                ("fd_tuned_graph_mesh_v1", V1, "FD", "FP64", "Persistent", "Schedule", True)]
    if args.c2_variant:
        variants = [v for v in variants if v[0] == args.c2_variant]
        if not variants:
            raise SystemExit(f"bench.py: no C2 variant named {args.c2_variant!r}")
    elif args.no_variants:
        variants = variants[:1]
    results = {}
    tmp = tempfile.mkdtemp(prefix="foort_c2_")
    # the autotuner's results are kept between runs of a configuration: the warm-up run tunes,
    # the timed runs reuse its choice
    os.environ["FOORT_TUNE_CACHE"] = os.path.join(tmp, "tune.txt")
    try:
        if args.warmup < 1:
            log("c2: with --warmup 0 the tuned variants include the autotuner's own trials")
        for name, mesh, chris, prec, sched, tune, graph in variants:
            cfg = os.path.join(tmp, f"{name}.cfg")
            with open(cfg, "w") as f:
                f.write(W.c2_config(mesh=mesh, christoffel=chris, precision=prec, schedule=sched,
                                    prefix=os.path.join(tmp, name), autotune=tune, use_graph=graph,
                                    tail_geodesics=args.tail_geodesics))
            walls, gpu_ms = [], []
            for i in range(args.warmup + args.steps):
                t = time.perf_counter()
                st = quiet(lambda: P.run(cfg, screen_output_level=2))
                dt = time.perf_counter() - t
                if i >= args.warmup:
                    walls.append(dt)
                    gpu_ms.append(st.device_ms)
                log(f"c2 {name} run {i}: {st.geodesics} pixels, {st.steps} steps, {dt:.2f} s wall, "
                    f"{st.device_ms / 1e3:.2f} s kernels")
            wall = sum(walls) / len(walls)
            results[name] = {"pixels": st.geodesics, "geodesic_steps": st.steps, "wall_s": wall,
                             "kernel_s": sum(gpu_ms) / len(gpu_ms) / 1e3, "pixels_per_s": st.geodesics / wall,
                             "geodesic_steps_per_s": st.steps / wall}
        first = results[variants[0][0]]
        line = {"metric": "pixels/s of a complete adaptive-mesh run", "value": first["pixels_per_s"],
                "unit": "pixels/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                "ms_per_step": first["wall_s"] * 1e3, "higher_is_better": True, "scaling": "weak",
                "vs_baseline": None, "dtype": "f64", "data": "synthetic",
                "config": {"workload": W.C2_TEXT, "path": variants[0][0]}, "variants": results}
        print(json.dumps(line), flush=True)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def run_c3(args):
    """C3: precision study.  Kerr a=0.998 at inclinations 17 and 85 deg, 2048x2048 uniform grid
    (otherwise as C1).  The mixed path (binary32 steps far from the hole and the poles) and the
    analytic binary64 path against the finite-difference binary64 path, which is bit for bit the
    CPU oracle: integer-diagnostic mismatch rates over all pixels and over the pixels none of
    whose four neighbours differ in the oracle image (the photon-ring boundary is chaotic), and
    relative errors of the final positions."""
    import numpy as np
    import foort_amd as F
    import foort_amd.abi as A
    from foort_amd import workloads as W
    side = args.side if args.side != 1024 else 2048
    n = side * side
    paths = [("fd_fp64", dict()),
             ("analytic_fp64", dict(christoffel=A.CHRISTOFFEL_ANALYTIC)),
             ("mixed", dict(christoffel=A.CHRISTOFFEL_ANALYTIC, precision=A.PRECISION_MIXED))]
    res = F.device_results(n, fields=("term", "quadrant", "equat_passes", "final_pos"))
    study = {}
    for inc in (17.0, 85.0):
        out = {}
        for name, kw in paths:
            with F.Context(W.c1(params=(0.998,), inclination_deg=inc, **kw)) as ctx:
                st = ctx.trace_device(side, 0, n, res)
            out[name] = {k: v.numpy() for k, v in res.items()}
            out[name]["stats"] = st
            log(f"c3 {inc:.0f} deg {name}: {st.device_ms / 1e3:.2f} s kernels, {st.steps} steps")
        ref = out["fd_fp64"]
        ints = np.stack([ref["term"], ref["quadrant"], ref["equat_passes"]], -1).reshape(side, side, 3)
        edge = np.zeros((side, side), bool)
        for ax in (0, 1):
            d = np.any(np.diff(ints, axis=ax) != 0, axis=-1)
            if ax == 0:
                edge[1:, :] |= d
                edge[:-1, :] |= d
            else:
                edge[:, 1:] |= d
                edge[:, :-1] |= d
        edge = edge.reshape(-1)
        entry = {"boundary_pixel_fraction": float(edge.mean())}
        for name in ("analytic_fp64", "mixed"):
            o = out[name]
            mism = (o["term"] != ref["term"]) | (o["quadrant"] != ref["quadrant"]) | \
                   (o["equat_passes"] != ref["equat_passes"])
            same = ~mism
            rel = np.abs(o["final_pos"] - ref["final_pos"]) / np.maximum(np.abs(ref["final_pos"]), 1e-300)
            relmax = rel.max(axis=1)[same]
            st = o["stats"]
            entry[name] = {
                "integer_diag_mismatch_rate": float(mism.mean()),
                "mismatch_rate_off_boundary": float(mism[~edge].mean()),
                "mismatched_pixels": int(mism.sum()),
                "mismatched_off_boundary": int(mism[~edge].sum()),
                "pos_rel_err_median": float(np.median(relmax)),
                "pos_rel_err_p99": float(np.quantile(relmax, 0.99)),
                "pos_rel_err_max": float(relmax.max()),
                "kernel_s": st.device_ms / 1e3,
                "pixels_per_s_kernel": n / (st.device_ms / 1e3),
                "speedup_vs_fd_kernel": out["fd_fp64"]["stats"].device_ms / st.device_ms,
            }
        entry["fd_fp64"] = {"kernel_s": out["fd_fp64"]["stats"].device_ms / 1e3,
                            "geodesic_steps": out["fd_fp64"]["stats"].steps}
        study[f"{inc:.0f}deg"] = entry
    mixed85 = study["85deg"]["mixed"]
    line = {"metric": "integer-diagnostic mismatch rate of the mixed path vs the FP64 oracle",
            "value": max(study[k]["mixed"]["integer_diag_mismatch_rate"] for k in study), "unit": "fraction",
            "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": mixed85["kernel_s"] * 1e3,
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
            "data": "synthetic",
            "config": {"workload": f"C3: Kerr a=0.998, 17 and 85 deg, {side}x{side}, otherwise as C1"},
            "study": study}
    print(json.dumps(line), flush=True)


def c4_measure(args):
    """The render server under a seeded open-loop request stream (Poisson arrivals): latency
    (submission to completion) percentiles, throughput, and whether the stream is sustained."""
    import numpy as np
    from foort_amd import program as P
    from foort_amd import workloads as W
    n = args.requests
    reqs = W.c4_requests(seed=args.seed, n=n, rate=args.rate)
    with P.Server(workers=args.workers, hbm_budget_bytes=args.hbm_budget << 20,
                  launch_us=args.launch_us or 2000) as srv:
        t0 = time.perf_counter()
        for k, (t, desc, mesh, _) in enumerate(reqs):
            wait = t - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            srv.submit(desc, mesh, k)
            if k % 250 == 249:
                log(f"c4: {k + 1} of {n} requests submitted")
        srv.drain()
        res = srv.results(cap=n + 16)
    assert len(res) == n and all(r.status == 0 for r in res), [r.status for r in res if r.status]
    sub = np.array([r.submit_s for r in res])
    lat = np.array([r.end_s - r.submit_s for r in res])
    wait = np.array([r.start_s - r.submit_s for r in res])
    span = max(r.end_s for r in res) - sub.min()
    pixels = sum(r.pixels for r in res)
    kernel_s = sum(r.kernel_ms for r in res) / 1e3
    out = {"workload": W.C4_TEXT.format(seed=args.seed, rate=args.rate), "requests": n, "workers": args.workers,
           "ring_bytes_per_group": args.hbm_budget << 20, "launch_us": args.launch_us or 2000,
           "latency_p50_s": float(np.quantile(lat, 0.5)), "latency_p99_s": float(np.quantile(lat, 0.99)),
           "latency_max_s": float(lat.max()), "queue_wait_p99_s": float(np.quantile(wait, 0.99)),
           "offered_rate_per_s": args.rate, "completed_rate_per_s": n / span,
           "pixels_per_s": pixels / span, "pixels": int(pixels), "geodesic_steps": int(sum(r.steps for r in res)),
           "gpu_kernel_s_shared_out": kernel_s, "span_s": span}
    # whether the stream is sustained: latency does not grow along the stream, and the requests
    # in flight when the last one arrives drain in about one request's latency
    end = np.array([r.end_s for r in res])
    order = np.argsort(sub)
    quarters = np.array_split(order, 4)
    out["sustained"] = {
        "arrival_span_s": float(sub.max() - sub.min()),
        "latency_p50_by_arrival_quarter_s": [float(np.quantile(lat[q], 0.5)) for q in quarters],
        "latency_p99_by_arrival_quarter_s": [float(np.quantile(lat[q], 0.99)) for q in quarters],
        "in_flight_at_last_arrival": int(((sub <= sub.max()) & (end > sub.max())).sum()),
        "drain_after_last_arrival_s": float(end.max() - sub.max()),
    }
    by_kind = {}
    for r in res:
        by_kind.setdefault(reqs[r.tag][3]["kind"], []).append(r)
    out["by_kind"] = {
        k: {"requests": len(v), "latency_p50_s": float(np.quantile([r.end_s - r.submit_s for r in v], 0.5)),
            "latency_p99_s": float(np.quantile([r.end_s - r.submit_s for r in v], 0.99)),
            "steps_per_pixel": sum(r.steps for r in v) / max(1, sum(r.pixels for r in v))}
        for k, v in sorted(by_kind.items())}
    log(f"c4: p50 {out['latency_p50_s']:.3f} s, p99 {out['latency_p99_s']:.3f} s, "
        f"{out['completed_rate_per_s']:.1f} requests/s completed of {args.rate:g}/s offered")
    return out


def run_c4(args):
    """C4 alone: the render server's line (value = latency p99)."""
    m = c4_measure(args)
    line = {"metric": "render-server request latency p99", "value": m["latency_p99_s"], "unit": "s",
            "n_gpus": 1, "steps": args.requests, "warmup": 0, "ms_per_step": m["latency_p50_s"] * 1e3,
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
            "data": "synthetic", "config": {"workload": m["workload"]}}
    line.update({k: v for k, v in m.items() if k != "workload"})
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["c1", "c2", "c3", "c4"], default="c1",
                    help="c1: the 1024^2 drop-in image (default); c2: the adaptive-mesh run; "
                         "c3: the precision study; c4: the render server")
    ap.add_argument("--requests", type=int, default=2000, help="c4: requests in the stream")
    ap.add_argument("--rate", type=float, default=40.0, help="c4: offered request rate (per s)")
    ap.add_argument("--seed", type=int, default=2024, help="c4: seed of the request stream")
    ap.add_argument("--workers", type=int, default=4, help="c4: host threads for mesh work")
    ap.add_argument("--hbm-budget", type=int, default=4096, help="c4: device ring size per request group in MiB")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--side", type=int, default=1024, help="image side (1024 is the workload)")
    ap.add_argument("--cpu-stride", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--no-server", action="store_true", help="c1: skip the render-server (C4) leg")
    ap.add_argument("--tail-geodesics", type=int, default=None,
                    help="c1, c2: unfinished geodesics handed to the four-lanes-per-geodesic "
                         "kernel (finite differences, persistent schedule; 4294967295 = never)")
    ap.add_argument("--c2-variant", default=None,
                    help="c2: run only the named variant (fd_tuned_graph, fd_oneshot, fd_persistent, ...)")
    ap.add_argument("--christoffel", choices=["fd", "analytic"], default="fd",
                    help="timed path: the reference's finite differences (default) or analytic")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--schedule", choices=["oneshot", "persistent"], default="oneshot")
    ap.add_argument("--steps-per-launch", type=int, default=0, help="persistent schedule budget (0 = default)")
    ap.add_argument("--block", type=int, default=0, help="one-shot block size (0 = default)")
    ap.add_argument("--refill", type=int, default=0, help="persistent schedule refill threshold (0 = default)")
    ap.add_argument("--interval", type=int, default=0, help="persistent schedule steps between refills (0 = default)")
    ap.add_argument("--launch-us", type=int, default=0, help="persistent schedule time per launch (0 = unbounded)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}); "
              f"pixels are independent, so the path would shard by interleaved pixel index, one process per GPU",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "c2":
        run_c2(args)
        return
    if args.workload == "c3":
        run_c3(args)
        return
    if args.workload == "c4":
        run_c4(args)
        return

    import foort_amd as F
    from foort_amd import workloads as W

    import foort_amd.abi as A
    side = args.side
    n = side * side
    path = dict(christoffel=A.CHRISTOFFEL_FD if args.christoffel == "fd" else A.CHRISTOFFEL_ANALYTIC,
                precision=A.PRECISION_FP64 if args.precision == "fp64" else A.PRECISION_MIXED,
                schedule=A.SCHEDULE_ONESHOT if args.schedule == "oneshot" else A.SCHEDULE_PERSISTENT,
                steps_per_launch=args.steps_per_launch, block_size=args.block, refill_lanes=args.refill,
                refill_interval=args.interval, launch_us=args.launch_us)
    if args.tail_geodesics is not None:
        path["tail_geodesics"] = args.tail_geodesics
    desc = W.c1(**path)
    ref_desc = W.c1()
    res = F.device_results(n)
    ctx = F.Context(desc)
    for i in range(args.warmup):
        st = ctx.trace_device(side, 0, n, res)
        log(f"warmup {i}: {st.device_ms:.1f} ms, {st.steps} steps")
    F.synchronize()
    kernel_ms, steps = [], []
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = ctx.trace_device(side, 0, n, res)  # waits for completion and reads the HIP events
        kernel_ms.append(st.device_ms)
        steps.append(st.steps)
        log(f"step {i}: {st.device_ms:.1f} ms, {st.steps} steps")
    F.synchronize()
    wall = time.perf_counter() - t0
    base = {k: v.numpy() for k, v in res.items()}
    ctx.close()

    value = args.steps * n / wall
    closed, fd = flops_per_step("kerr")
    avg_ms = sum(kernel_ms) / len(kernel_ms)
    kname = {"fd": "RhsFD", "analytic": "RhsAn64"}[args.christoffel] if args.precision == "fp64" else "RhsMixed"
    kernel = f"trace_{args.schedule}<{kname}<Kerr>, RK4>"
    steps_per_launch = sum(steps) / len(steps)
    achieved = closed["rk4_step_flops"] * steps_per_launch / (avg_ms * 1e-3) / 1e12
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "pixels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": "C1: " + W.C1_TEXT, "pixels": n,
                   "christoffel": "finite differences (reference)" if args.christoffel == "fd" else "analytic",
                   "precision": args.precision, "schedule": args.schedule,
                   "parity": "bit-identical to the CPU oracle" if args.christoffel == "fd"
                   else "statistical (see variants)"},
        "geodesic_steps_per_s": sum(steps) / wall,
        "roofline": {
            "bound": "valu_fp64",
            "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS,
            "traffic": (lambda b: b * n if b is not None else None)(
                hbm_bytes_per_pixel(args.christoffel, args.precision, args.schedule)),
            "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE per pixel of this path (calibrated on a device "
                              "copy, the 1024x1024 image) x the launch's pixels: profiles/hbm_traffic_kerr_rk4_c1.md",
            "kernel": kernel,
            "flops_per_step": closed["rk4_step_flops"],
            "flops_per_step_formulation": closed["formulation"],
            "steps_per_launch": steps_per_launch,
            "avg_launch_ms": avg_ms,
            "fd_formulation_flops_per_step": fd["rk4_step_flops"],
            "fd_formulation_transcendentals_per_step": fd["rk4_step_transcendental"],
        },
    }

    if not args.no_variants:
        import foort_amd.abi as A
        variants = {}
        for name, kw in [("fd_persistent", dict(schedule=A.SCHEDULE_PERSISTENT)),
                         ("analytic_fp64_oneshot", dict(christoffel=A.CHRISTOFFEL_ANALYTIC)),
                         ("analytic_fp64_persistent", dict(christoffel=A.CHRISTOFFEL_ANALYTIC,
                                                           schedule=A.SCHEDULE_PERSISTENT)),
                         ("mixed_oneshot", dict(christoffel=A.CHRISTOFFEL_ANALYTIC, precision=A.PRECISION_MIXED)),
                         ("mixed_persistent", dict(christoffel=A.CHRISTOFFEL_ANALYTIC, precision=A.PRECISION_MIXED,
                                                   schedule=A.SCHEDULE_PERSISTENT))]:
            with F.Context(W.c1(**kw)) as vctx:
                tuned = None
                if name == "mixed_persistent":
                    # the north-star path with the options the project's autotuner picks for this
                    # image (steps between refills, refill threshold; its trials are not timed)
                    chosen, _ = vctx.autotune(side, count=n)
                    tuned = {"refill_interval": int(chosen.refill_interval), "refill_lanes": int(chosen.refill_lanes)}
                vctx.trace_device(side, 0, n, res)  # warm-up
                F.synchronize()
                t = time.perf_counter()
                st = vctx.trace_device(side, 0, n, res)
                F.synchronize()
                dt = time.perf_counter() - t
            out = {k: v.numpy() for k, v in res.items()}
            c = "analytic" if kw.get("christoffel") == A.CHRISTOFFEL_ANALYTIC else "fd"
            p = "mixed" if kw.get("precision") == A.PRECISION_MIXED else "fp64"
            s = "persistent" if kw.get("schedule") == A.SCHEDULE_PERSISTENT else "oneshot"
            variants[name] = {"pixels_per_s": n / dt, "geodesic_steps_per_s": st.steps / dt,
                              "kernel_ms": st.device_ms, "launches": st.launches,
                              "hbm_bytes_per_pixel": hbm_bytes_per_pixel(c, p, s),
                              "vs_fd_oneshot": agreement(base, out)}
            if tuned:
                variants[name]["autotuned"] = tuned
            log(f"{name}: {n / dt:.3e} px/s")
        line["variants"] = variants

    if not args.no_variants:
        # the north-star ratio: the mixed-precision persistent path over the plain FP64 path
        line["mixed_persistent_vs_fd_fp64"] = line["variants"]["mixed_persistent"]["pixels_per_s"] / value

    if not args.no_server:
        # the headline metric's latency half: the render server under the C4 request stream
        log("render server (C4 stream) ...")
        srv = c4_measure(args)
        line["latency_p50_s"] = srv["latency_p50_s"]
        line["latency_p99_s"] = srv["latency_p99_s"]
        line["render_server"] = srv

    if not args.no_cpu:
        log("cpu baseline ...")
        line["cpu_baseline"] = cpu_baseline(ref_desc, side, args.cpu_stride)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
