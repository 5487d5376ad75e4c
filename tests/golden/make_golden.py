"""Write the golden fixtures of tests/golden/ (run in the container that holds
the reference tree; the outputs are committed, the reference never travels).

1. reference_outputs.npz -- the integer diagnostics the reference itself wrote
   to FOORT/Output/{Test,debug}_{FourColorScreen,EquatPasses}.dat (250 x 250
   SimpleSquareMesh runs of FOORT/Test.cfg and of the run described in the
   debug files' first line), as int8 arrays indexed [row, column].
2. ref_states.npz -- final states computed by the reference's own sources
   (oracle/_ref/libfoort_ref.so, IEEE flags) for the parity cases in
   tests/cases.py: termination, diagnostics, final position, velocity,
   affine parameter and step count per pixel.
3. sample_outputs.npz -- the pixels and integer values of the reference's
   sample runs (FOORT/Output/sampleoutput_*).

Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
REF_OUT = "/root/reference/FOORT/Output"


def read_dat(path, side=250):
    with open(path) as f:
        header = f.readline().rstrip("\n")
        a = np.full((side, side), -128, np.int16)
        for line in f:
            row, col, val = line.split()
            a[int(row), int(col)] = int(val)
    assert (a != -128).all(), path
    return header, a.astype(np.int8)


def reference_outputs():
    out = {}
    for run in ("Test", "debug"):
        for diag, key in (("FourColorScreen", "fcs"), ("EquatPasses", "ep")):
            header, arr = read_dat(os.path.join(REF_OUT, f"{run}_{diag}.dat"))
            out[f"{run}_{key}"] = arr
            out[f"{run}_{key}_header"] = np.array(header)
    np.savez_compressed(os.path.join(HERE, "reference_outputs.npz"), **out)


def ref_states():
    import oracle as O
    from tests.cases import NO_REFERENCE_COUNTERPART, PARITY_CASES
    out = {}
    for name, (mk, side, px) in sorted(PARITY_CASES.items()):
        if name in NO_REFERENCE_COUNTERPART:
            continue
        desc = mk()
        r = O.trace("ref", desc, side) if px is None else O.trace("ref", desc, side, screen_index=px)
        for k, v in r.items():
            out[f"{name}/{k}"] = v
    np.savez_compressed(os.path.join(HERE, "ref_states.npz"), **out)


# the reference's sample runs (FOORT/Output/sampleoutput_*): Kerr a = 0.5, r coordinate,
# 15 x 15 screen, camera at 90 deg (simple and adaptive mesh) and at 0.296697 rad (17 deg,
# adaptive, "rings"); name -> (file stem, diagnostics, side of the finest grid)
SAMPLES = {
    "simplemesh": ("sampleoutput_simplemesh_221202-125440", ("FourColorScreen",), 100),
    "adaptivemesh": ("sampleoutput_adaptivemesh_221202-125808", ("FourColorScreen",), (100 - 1) * 8 + 1),
    "rings": ("sampleoutput_adaptivemesh_rings_221202-140701", ("FourColorScreen", "EquatPasses"), (141 - 1) * 8 + 1),
}


def sample_outputs():
    """sample_outputs.npz -- the pixels (row, column on the finest grid) and integer values of
    the reference's sample output files, in file order, with their first lines."""
    out = {}
    for name, (stem, diags, side) in SAMPLES.items():
        rows = cols = None
        for diag in diags:
            with open(os.path.join(REF_OUT, f"{stem}_{diag}.dat")) as f:
                header = f.readline().rstrip("\n")
                rc, vals = [], []
                for line in f:
                    r, c, v = line.split()
                    rc.append((int(r), int(c)))
                    vals.append(int(v))
            rc = np.array(rc, np.uint16)
            assert rc.max() < side, name
            if rows is None:
                rows, cols = rc[:, 0], rc[:, 1]
            assert (rows == rc[:, 0]).all() and (cols == rc[:, 1]).all(), name
            out[f"{name}_{diag}"] = np.array(vals, np.int8)
            out[f"{name}_{diag}_header"] = np.array(header)
        out[f"{name}_rows"], out[f"{name}_cols"] = rows, cols
        out[f"{name}_side"] = np.array(side)
    np.savez_compressed(os.path.join(HERE, "sample_outputs.npz"), **out)


if __name__ == "__main__":
    reference_outputs()
    ref_states()
    sample_outputs()
    print("written", os.listdir(HERE))
