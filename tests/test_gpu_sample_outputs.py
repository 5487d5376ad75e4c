"""The reference's sample runs (FOORT/Output/sampleoutput_*, tests/golden/sample_outputs.npz):
Kerr a = 0.5 in the r coordinate, 15 x 15 screen, horizon 1.01 r_h, boundary 1000, time-out
1e6 steps; a 100 x 100 simple mesh and an adaptive mesh refined to a 793 x 793 grid at 90 deg,
and an adaptive mesh refined to 1121 x 1121 at 17 deg ("rings", camera angle 17 x 3.1415/180
as its first line prints it).

These files were written by an older version of the reference than the sources here.  Its
screen has the opposite handedness: column c of a file is column side - 1 - c here, and
FourColorScreen's two halves in phi are swapped (1 <-> 2, 3 <-> 4).  With that mirror:

* the 100 x 100 and adaptive 90 deg files agree on every pixel except pixels with a neighbour
  of another value (38 of 10,000 and 2,210 of 33,817, all on the boundaries);
* the 17 deg file agrees on every pixel inside 4 M and outside 8 M of the screen centre
  (13,141 pixels); in the photon-ring band between, where the adaptive mesh puts 81% of the
  file's pixels on a 1121 x 1121 sampling of the chaotic region, the older version's values
  differ on 22% of them, boundary or not.  That band is not pinned here;
  the current sources' own shipped files (FOORT/Output/Test_*.dat, debug_*.dat) are reproduced
  byte for byte (test_gpu_program.py).
"""
import math
import os

import numpy as np
import pytest

from foort_amd import abi as A

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SWAP_PHI = np.array([0, 2, 1, 4, 3])


def _boundary(a):
    b = np.zeros(a.shape, bool)
    d = a[1:] != a[:-1]
    b[1:] |= d
    b[:-1] |= d
    d = a[:, 1:] != a[:, :-1]
    b[:, 1:] |= d
    b[:, :-1] |= d
    return b


def _image(theta, side):
    import foort_amd
    desc = A.make_desc(A.METRIC_KERR, (0.5,), rlog=False, theta=theta, screen=(15.0, 15.0), max_steps=1000000,
                       horizon_eps=0.01, diags=A.DIAGFLAG_FOURCOLORSCREEN)
    with foort_amd.Context(desc) as ctx:
        q = ctx.trace_grid(side)["quadrant"].reshape(side, side)
    return SWAP_PHI[q[:, ::-1]]  # into the older version's screen


@pytest.mark.parametrize("name", ["simplemesh", "adaptivemesh"])
def test_sample_outputs_at_90_degrees(name):
    d = np.load(os.path.join(GOLD, "sample_outputs.npz"))
    side = int(d[f"{name}_side"])
    print(f"[sample] {name} {side}x{side}", flush=True)
    img = _image(math.pi / 2, side)
    r, c = d[f"{name}_rows"].astype(int), d[f"{name}_cols"].astype(int)
    mis = img[r, c] != d[f"{name}_FourColorScreen"]
    assert not (mis & ~_boundary(img)[r, c]).any()
    assert mis.mean() < {"simplemesh": 0.01, "adaptivemesh": 0.1}[name]


def test_sample_output_rings_at_17_degrees_outside_the_photon_ring_band():
    d = np.load(os.path.join(GOLD, "sample_outputs.npz"))
    side = int(d["rings_side"])
    print(f"[sample] rings {side}x{side}", flush=True)
    img = _image(17 * 3.1415 / 180, side)
    r, c = d["rings_rows"].astype(int), d["rings_cols"].astype(int)
    mid = (side - 1) / 2
    rad = np.hypot(r - mid, c - mid) / mid * 7.5  # distance from the screen centre in M
    band = (rad >= 4.0) & (rad < 8.0)
    mis = img[r, c] != d["rings_FourColorScreen"]
    assert not (mis & ~band).any()
    assert (~band).sum() > 13000
