"""Edge cases of the three entry points: degenerate grids, grids with no
point inside the images, images smaller than a window, a single point,
all-invalid and one-cell cull grids.  Outputs are checked against the
oracle (same inputs) or against the reference's no-data encoding
(offsets -2e9, type 0 for strack, type 2 for strackw)."""
import numpy as np
import pytest

from specklesource_amd import abi, configs, track
from tests.helpers import cull_params, make_cull_grid, small_params
from tests.test_oracle_strackw import stw_params

pytestmark = pytest.mark.gpu

NODATA = np.float32(-2e9)


def _scene(n=256, seed=4):
    s = configs._spec(n, n, seed, dr0=0.7, da0=-0.4, gmin=0.8, gmax=0.8)
    return track.synth_host(s)


def test_zero_sized_grid_is_an_argument_error(device):
    a, b = _scene()
    pid = device.upload_pair(a, b)
    try:
        p = small_params(32, 256, 256, nR=0)
        with pytest.raises(RuntimeError):
            device.speckle_track(pid, p)
        with pytest.raises(RuntimeError):
            device.corr_track_fast(pid, stw_params(64, 256, 256, edge=8, nA=0))
    finally:
        device.release_pair(pid)


@pytest.mark.parametrize("w", [16, 64, 128])
def test_grid_outside_images_gives_no_data(device, oracle32, w):
    """Every grid point's window falls outside image 2 (a 5000 px initial
    offset): nothing is tracked, every point keeps the no-data encoding."""
    a, b = _scene()
    p = small_params(w, 256, 256, step=w // 2, rShiftPoly=[5000.0, 0, 0])
    pid = device.upload_pair(a, b)
    try:
        g = device.speckle_track(pid, p)
    finally:
        device.release_pair(pid)
    o = oracle32.speckle_track(p, a, b)
    assert (g.type == 0).all() and (g.offR == NODATA).all() and (g.offA == NODATA).all()
    assert g.nComplex == g.nAmp == g.nAmpL == 0
    assert np.array_equal(g.type, o.type) and np.array_equal(g.offR, o.offR)


def test_single_point_and_image_smaller_than_window(device, oracle32):
    a, b = _scene(96)
    p = small_params(32, 96, 96, nR=1, nA=1, rStart=48, aStart=48)
    pid = device.upload_pair(a, b)
    try:
        g = device.speckle_track(pid, p)
        # 128 x 128 windows on a 96 x 96 pair: every window leaves the image
        g2 = device.speckle_track(pid, small_params(128, 96, 96, nR=2, nA=2, rStart=40, aStart=40, deltaR=8,
                                                    deltaA=8))
    finally:
        device.release_pair(pid)
    o = oracle32.speckle_track(p, a, b)
    assert g.type.shape == (1, 1)
    assert np.array_equal(g.type, o.type) and np.array_equal(g.offR, o.offR) and np.array_equal(g.offA, o.offA)
    assert (g2.type == 0).all() and (g2.offR == NODATA).all()


def test_strackw_with_no_listed_point(device, oracle32):
    a, b = _scene(512)
    p = stw_params(64, 512, 512, edge=8, navg=1, step=32, rShiftPoly=[9000.0, 0, 0])
    pid = device.upload_pair(a, b)
    try:
        g = device.corr_track_fast(pid, p)
    finally:
        device.release_pair(pid)
    o = oracle32.corr_track_fast(p, a, b)
    assert g.nAmp == 0 and g.nFail == 0
    assert np.array_equal(g.type, o.type) and np.array_equal(g.offR, o.offR)
    assert (g.offR == NODATA).all()


def test_cull_all_invalid_and_single_cell(device, oracle32):
    offR, offA, corr, typ = make_cull_grid(23, 31, seed=9, outlier_frac=0.1)
    offR[:] = NODATA
    offA[:] = NODATA
    for shape in ((23, 31), (1, 1), (1, 40), (40, 1)):
        r, a_, c, t = (x[:shape[0], :shape[1]].copy() for x in (offR, offA, corr, typ))
        if shape == (1, 1):
            r[:] = 1.5
            a_[:] = -0.5
        gc, ga = cull_params(r, a_, c, t, islandThresh=20)
        oc, oa = cull_params(r, a_, c, t, islandThresh=20)
        device.cull("st_cull_run", gc)
        oracle32.cull("st_cull_run", oc)
        for k in ("offR", "offA", "type", "offRS", "offAS"):
            assert np.array_equal(ga[k], oa[k]), (shape, k)
