"""foort_gpu_configure on a live context.  Another run description (spin, camera, screen)
replaces the constants the kernels read; the persistent schedule's refills take the camera
frame from the context's device copy of them, which configure rewrites.  Every batch, before
and after each change, is checked bit for bit against the CPU oracle."""
import pytest

import foort_amd as F
import foort_amd.abi as A
import oracle as O
from tests.cases import FIELDS, bitwise_equal, c0


@pytest.mark.gpu
@pytest.mark.parametrize("schedule", [A.SCHEDULE_PERSISTENT, A.SCHEDULE_ONESHOT])
def test_configure_with_another_camera_matches_oracle(schedule):
    first = c0(params=(0.5,), rlog=True, inclination_deg=80.0, max_steps=5000, schedule=schedule)
    second = c0(params=(0.9,), rlog=True, inclination_deg=35.0, screen=(20.0, 12.0), center=(2.0, -1.0),
                max_steps=5000, schedule=schedule)
    side = 12
    runs = []
    with F.Context(first) as ctx:
        for desc in (first, second, first):
            ctx.configure(desc)
            runs.append((desc, ctx.trace_grid(side)))
    for n, (desc, g) in enumerate(runs):
        o = O.trace("fm", desc, side)
        bad = [k for k in FIELDS if not bitwise_equal(g[k], o[k])]
        assert not bad, f"batch {n}: GPU and oracle differ in {bad}"
