"""st_cull_run_noalloc, the cull chain that allocates no device memory,
against st_cull_run through the C ABI: identical grids on two calls in a
row, the second on a smaller grid (the shared scratch array keeps the
first call's contents), and a clear refusal above its size limit."""
import numpy as np
import pytest

from tests.helpers import cull_params, make_cull_grid

pytestmark = pytest.mark.gpu


def _run(device, name, grids, **kw):
    cp, arrs = cull_params(*grids, **kw)
    device.cull(name, cp)
    return arrs


def test_noalloc_cull_equals_cull_run_on_two_calls_second_smaller(device):
    big = make_cull_grid(571, 821, seed=21)  # the full-scene grid: 468 791 points
    small = make_cull_grid(60, 75, seed=22)
    for grids, kw in ((big, dict(islandThresh=20)), (small, dict(bR=5, bA=5, islandThresh=3))):
        want = _run(device, "st_cull_run", grids, **kw)
        got = _run(device, "st_cull_run_noalloc", grids, **kw)
        for f in want:
            assert np.array_equal(got[f], want[f]), f
        # the chain culled something (culled offsets become -2e9)
        assert (got["offR"] != grids[0]).any()


def test_noalloc_cull_refuses_grids_above_its_limit(device):
    grids = make_cull_grid(1024, 513, seed=1)  # 525 312 points > 524 288
    with pytest.raises(RuntimeError, match="no-allocation cull"):
        _run(device, "st_cull_run_noalloc", grids)
