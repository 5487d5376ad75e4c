"""Per-path HBM traffic of the integration kernels, from rocprofv3 counter runs.

bench.py reports `roofline.traffic` (HBM bytes per launch) from foort_amd/hbm_traffic.json,
which this script writes from two counter runs per path (FETCH_SIZE, then WRITE_SIZE, in KiB
per dispatch) on the benchmark's own 1024x1024 Kerr image, each counter scaled by what it reads
for a plain device copy of 256 MiB.  The commands that produce the runs (repo root; <c> <p> <s>
= Christoffel, precision, schedule of each path):

    rocprofv3 --pmc FETCH_SIZE -f csv -d <dir>/traffic_<c>_<p>_<s>_FETCH_SIZE -o r -- \\
        python3 -u bench.py --steps 1 --warmup 0 --no-variants --no-cpu --no-server \\
        --christoffel <c> --precision <p> --schedule <s>
    (the same with WRITE_SIZE)
    rocprofv3 --pmc FETCH_SIZE -f csv -d <dir>/copy_FETCH_SIZE -o r -- python3 -c "import torch; \\
        a=torch.ones(256<<20,dtype=torch.uint8,device='cuda'); b=torch.empty_like(a); b.copy_(a)"
    (the same with WRITE_SIZE)

    python3 foort_amd/counters.py <dir>      # writes foort_amd/hbm_traffic.json
"""
import csv
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "hbm_traffic.json")
PATHS = [("fd", "fp64", "oneshot"), ("fd", "fp64", "persistent"), ("analytic", "fp64", "oneshot"),
         ("analytic", "fp64", "persistent"), ("analytic", "mixed", "oneshot"), ("analytic", "mixed", "persistent")]
COPY_BYTES = 256 << 20
PIXELS = 1024 * 1024


def _dispatches(path):
    with open(path) as f:
        return list(csv.DictReader(f))


def _sum(rows, counter, select):
    return sum(float(r["Counter_Value"]) for r in rows if r["Counter_Name"] == counter and select(r["Kernel_Name"]))


def build(run_dir):
    """The traffic table from the counter runs in run_dir."""
    def csv_of(tag):
        return os.path.join(run_dir, tag, "r_counter_collection.csv")

    copy = {c: _sum(_dispatches(csv_of(f"copy_{c}")), c, lambda n: "copyBuffer" in n)
            for c in ("FETCH_SIZE", "WRITE_SIZE")}
    factor = {c: COPY_BYTES / 1024.0 / v for c, v in copy.items()}
    out = {"workload": "Kerr a=0.99, 85 deg, 1024x1024 SimpleSquareMesh, RK4, FP64 state (C1)",
           "pixels": PIXELS, "copy_kib": copy, "factor": factor, "paths": {}}
    for c, p, s in PATHS:
        entry = {"kernels": set()}
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            rows = _dispatches(csv_of(f"traffic_{c}_{p}_{s}_{counter}"))
            kib = _sum(rows, counter, lambda n: "foort::trace_" in n)
            entry["kernels"] |= {r["Kernel_Name"].split("(")[0].replace("void ", "") for r in rows
                                 if "foort::trace_" in r["Kernel_Name"]}
            entry[counter.lower() + "_kib"] = kib
            entry[counter.lower() + "_bytes"] = kib * 1024.0 * factor[counter]
        entry["kernels"] = sorted(entry["kernels"])
        entry["bytes_per_pixel"] = (entry["fetch_size_bytes"] + entry["write_size_bytes"]) / PIXELS
        out["paths"][f"{c}/{p}/{s}"] = entry
    return out


def load():
    """The table bench.py reads (None if it has not been written)."""
    try:
        with open(TABLE) as f:
            return json.load(f)
    except OSError:
        return None


def main(argv):
    if len(argv) != 2:
        raise SystemExit("usage: counters.py <directory of rocprofv3 counter runs>")
    t = build(argv[1])
    with open(TABLE, "w") as f:
        json.dump(t, f, indent=1)
    for k, v in t["paths"].items():
        print(f"{k}: {v['bytes_per_pixel']:.1f} B/pixel ({', '.join(v['kernels'])})")


if __name__ == "__main__":
    main(sys.argv)
