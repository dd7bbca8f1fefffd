"""GPU parity of the Cullls stages (Cullls/cullLSData.c, cullLSStats.c,
cullLSSmooth.c, cullLSIslands.c) against the oracle, through the C ABI.

Culled grids, the attempt statistics and island removal are compared bit
for bit; sigma and the smoothed offsets to the same 1-ulp-level tolerance
as the cullst stages (both sides run the same double-precision sums)."""
import numpy as np
import pytest

from specklesource_amd import track
from tests.test_gpu_cull import _close
from tests.test_oracle_cullls import LSKW, NODATA, SLOW, _ls_grid

pytestmark = pytest.mark.gpu


def _pair(X, Y, typ, **kw):
    k = dict(LSKW)
    k.update(kw)
    return track.cull_ls_params(X, Y, typ, **k), track.cull_ls_params(X, Y, typ, **k)


@pytest.mark.parametrize("shape,seed", [((40, 50), 1), ((97, 61), 2), ((16, 300), 3)])
def test_ls_pass_exact(device, oracle32, shape, seed):
    X, Y, typ = _ls_grid(*shape, seed=seed, outlier_frac=0.15)
    (gc, ga), (oc, oa) = _pair(X, Y, typ)
    for _ in range(3):
        device.cull("st_cull_ls_data", gc)
        oracle32.cull("st_cull_ls_data", oc)
        assert gc.nCulled == oc.nCulled
        assert np.array_equal(ga["X"], oa["X"]) and np.array_equal(ga["Y"], oa["Y"])
    assert (gc.nMatch, gc.nAttempt) == (oc.nMatch, oc.nAttempt)


@pytest.mark.parametrize("iters", ["48", "3", "0"])
def test_ls_schedules(device, oracle32, monkeypatch, iters):
    """A noisy ramp with a slow band: fixpoint, capped fixpoint + wavefront,
    and wavefront only all give the sequential result."""
    monkeypatch.setenv("SPECKLE_AMD_CULL_ITERS", iters)
    ny, nx = 30, 200
    rng = np.random.default_rng(11)
    jj = np.arange(nx)[None, :].repeat(ny, 0)
    X = (0.1 * jj + 0.35 * rng.standard_normal((ny, nx))).astype(np.float32)
    Y = (0.35 * rng.standard_normal((ny, nx))).astype(np.float32)
    typ = np.ones((ny, nx), np.uint8)
    typ[:, 60:140] |= SLOW
    (gc, ga), (oc, oa) = _pair(X, Y, typ, maxX=0.4, maxY=0.5, deltaT=40.0)
    device.cull("st_cull_ls_data", gc)
    oracle32.cull("st_cull_ls_data", oc)
    assert gc.nCulled == oc.nCulled
    assert np.array_equal(ga["X"], oa["X"]) and np.array_equal(ga["Y"], oa["Y"])


def test_ls_nodata_value(device, oracle32):
    X, Y, typ = _ls_grid(40, 40, seed=6)
    Y[10:14, 10:30] = -1000.0
    for nd in (-1000.0, -1001.5, -2e9):
        (gc, ga), (oc, oa) = _pair(X, Y, typ, no_data=nd)
        device.cull("st_cull_ls_data", gc)
        oracle32.cull("st_cull_ls_data", oc)
        assert gc.nCulled == oc.nCulled, nd
        assert np.array_equal(ga["X"], oa["X"]) and np.array_equal(ga["Y"], oa["Y"]), nd


def test_ls_stats_smooth(device, oracle32):
    X, Y, typ = _ls_grid(60, 70, seed=9)
    (gc, ga), (oc, oa) = _pair(X, Y, typ)
    for _ in range(3):
        oracle32.cull("st_cull_ls_data", oc)
    ga["X"][:] = oa["X"]
    ga["Y"][:] = oa["Y"]
    device.cull("st_cull_ls_stats", gc)
    oracle32.cull("st_cull_ls_stats", oc)
    assert np.array_equal(ga["sigmaX"] < -1e9, oa["sigmaX"] < -1e9)
    for k in ("sigmaX", "sigmaY"):
        ok, err = _close(ga[k], oa[k])
        assert ok, (k, err)
    for s in ((3, 3), (1, 5), (4, 2)):
        ga["sigmaX"][:] = oa["sigmaX"]
        ga["sigmaY"][:] = oa["sigmaY"]
        sx, sy = s
        gc.sX, gc.sY, oc.sX, oc.sY = sx, sy, sx, sy
        device.cull("st_cull_ls_smooth", gc)
        oracle32.cull("st_cull_ls_smooth", oc)
        assert np.array_equal(ga["XS"] < -1e9, oa["XS"] < -1e9), s
        for k in ("XS", "YS", "sigmaX", "sigmaY"):
            ok, err = _close(ga[k], oa[k])
            assert ok, (s, k, err)
        oracle32.cull("st_cull_ls_stats", oc)  # restore the unsmoothed sigmas for the next window


@pytest.mark.parametrize("thresh", [1, 4])
def test_ls_islands_exact(device, oracle32, thresh):
    rng = np.random.default_rng(30 + thresh)
    valid = rng.random((70, 90)) < 0.5
    XS = np.where(valid, 1.0, NODATA).astype(np.float32)
    (gc, ga), (oc, oa) = _pair(XS, XS, np.ones(XS.shape, np.uint8), islandThresh=thresh)
    for a in (ga, oa):
        a["XS"][:] = XS
        a["YS"][:] = XS
    device.cull("st_cull_ls_islands", gc)
    oracle32.cull("st_cull_ls_islands", oc)
    assert np.array_equal(ga["XS"], oa["XS"]) and np.array_equal(ga["YS"], oa["YS"])


def test_ls_run(device, oracle32):
    X, Y, typ = _ls_grid(120, 90, seed=13, outlier_frac=0.12)
    (gc, ga), (oc, oa) = _pair(X, Y, typ, islandThresh=3)
    device.cull("st_cull_ls_run", gc)
    oracle32.cull("st_cull_ls_run", oc)
    assert (gc.nMatch, gc.nAttempt) == (oc.nMatch, oc.nAttempt)
    assert np.array_equal(ga["X"], oa["X"]) and np.array_equal(ga["Y"], oa["Y"])
    assert np.array_equal(ga["XS"] < -1e9, oa["XS"] < -1e9)
    for k in ("XS", "YS", "sigmaX", "sigmaY"):
        ok, err = _close(ga[k], oa[k])
        assert ok, (k, err)
