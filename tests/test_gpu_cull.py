"""GPU parity of the cull stages (Cullst/cullSTData.c, cullStats.c,
cullSmooth.c, cullIslands.c) against the oracle, through the C ABI.

Cull decisions, culled offsets, types and island removal are compared bit
for bit.  sigma and the smoothed offsets follow the same operation order in
double precision as the oracle, so they are compared to 1 ulp-level
tolerances written below."""
import numpy as np
import pytest

from tests.helpers import cull_params, make_cull_grid

pytestmark = pytest.mark.gpu

# sigma / smoothed offsets: relative tolerance (the GPU and the oracle run
# the same double-precision sums; only libm-level rounding may differ)
RTOL = 2e-6


def _pair(offR, offA, corr, typ, mask=None, **kw):
    g = cull_params(offR, offA, corr, typ, mask=mask, **kw)
    o = cull_params(offR, offA, corr, typ, mask=mask, **kw)
    return g, o


def _close(a, b, rtol=RTOL):
    ok = np.isclose(a, b, rtol=rtol, atol=1e-6) | ((a < -1e9) & (b < -1e9))
    return bool(ok.all()), float(np.max(np.abs(a.astype(np.float64) - b)[~((a < -1e9) & (b < -1e9))], initial=0))


@pytest.mark.parametrize("shape,seed,frac", [((40, 50), 1, 0.05), ((97, 61), 2, 0.3), ((16, 300), 3, 0.15)])
def test_cull_pass_exact(device, oracle32, shape, seed, frac):
    offR, offA, corr, typ = make_cull_grid(*shape, seed=seed, outlier_frac=frac)
    (gc, ga), (oc, oa) = _pair(offR, offA, corr, typ)
    for _ in range(3):
        device.cull("st_cull_st_data", gc)
        oracle32.cull("st_cull_st_data", oc)
        assert gc.nCulled == oc.nCulled
        assert np.array_equal(ga["offR"], oa["offR"]) and np.array_equal(ga["offA"], oa["offA"])
        assert np.array_equal(ga["type"], oa["type"])


def test_cull_step_edge_in_place_order(device, oracle32):
    """A noisy step edge, where the in-place raster order changes the
    result against a parallel pass."""
    rng = np.random.default_rng(6)
    offR = np.zeros((24, 24), np.float32)
    offA = np.zeros((24, 24), np.float32)
    offR[:, 12:] = 1.0
    offR += (0.3 * rng.standard_normal(offR.shape)).astype(np.float32)
    (gc, ga), (oc, oa) = _pair(offR, offA, np.ones_like(offR), np.ones(offR.shape, np.uint8))
    device.cull("st_cull_st_data", gc)
    oracle32.cull("st_cull_st_data", oc)
    assert np.array_equal(ga["offR"], oa["offR"])


@pytest.mark.parametrize("iters", ["48", "3", "0"])
def test_cull_schedules(device, oracle32, monkeypatch, iters):
    """A noisy ramp needs several fixpoint iterations; capping them at 3 or
    0 hands the pass (or all of it) to the wavefront kernel, and every
    schedule must give the sequential result."""
    monkeypatch.setenv("SPECKLE_AMD_CULL_ITERS", iters)
    nA, nR = 30, 200
    rng = np.random.default_rng(11)
    jj = np.arange(nR)[None, :].repeat(nA, 0)
    offR = (0.1 * jj + 0.35 * rng.standard_normal((nA, nR))).astype(np.float32)
    offA = (0.35 * rng.standard_normal((nA, nR))).astype(np.float32)
    (gc, ga), (oc, oa) = _pair(offR, offA, np.ones_like(offR), np.ones(offR.shape, np.uint8),
                               maxR=0.4, maxA=0.5)
    device.cull("st_cull_st_data", gc)
    oracle32.cull("st_cull_st_data", oc)
    assert gc.nCulled == oc.nCulled
    assert np.array_equal(ga["offR"], oa["offR"]) and np.array_equal(ga["offA"], oa["offA"])


def test_cull_prepass_mask_corr_singlemt(device, oracle32):
    offR, offA, corr, typ = make_cull_grid(50, 40, seed=4)
    mask = np.random.default_rng(5).integers(0, 8, offR.shape).astype(np.uint8)
    for kw in (dict(corrThresh=0.3), dict(singleMT=2), dict(corrThresh=0.1, singleMT=1)):
        (gc, ga), (oc, oa) = _pair(offR, offA, corr, typ, mask=mask, **kw)
        device.cull("st_cull_st_data", gc)
        oracle32.cull("st_cull_st_data", oc)
        assert gc.nCulled == oc.nCulled, kw
        assert np.array_equal(ga["offR"], oa["offR"]), kw
        assert np.array_equal(ga["type"], oa["type"]), kw


@pytest.mark.parametrize("box", [(3, 3), (5, 11), (15, 15), (4, 6), (21, 21), (17, 31), (45, 45), (44, 20)])
def test_cull_box_sizes(device, oracle32, box):
    """Boxes up to 45 x 45 (the per-wave LDS lists are sized per call; an
    even side b spans b + 1 cells, as in cullSTData.c:73-80)."""
    offR, offA, corr, typ = make_cull_grid(45, 52, seed=8, outlier_frac=0.1)
    bR, bA = box
    (gc, ga), (oc, oa) = _pair(offR, offA, corr, typ, bR=bR, bA=bA, nGood=min(60, max(1, bR * bA // 5)))
    device.cull("st_cull_st_data", gc)
    oracle32.cull("st_cull_st_data", oc)
    assert gc.nCulled == oc.nCulled
    assert np.array_equal(ga["offR"], oa["offR"]) and np.array_equal(ga["offA"], oa["offA"])


@pytest.mark.parametrize("iters", ["48", "0"])
def test_cull_large_box_schedules(device, oracle32, monkeypatch, iters):
    """A 41 x 41 box through both schedules: the fixpoint kernel, and the
    sequential wavefront kernel, which runs fewer waves when the box lists
    would not fit 16 of them in the workgroup's LDS."""
    monkeypatch.setenv("SPECKLE_AMD_CULL_ITERS", iters)
    offR, offA, corr, typ = make_cull_grid(30, 60, seed=12, outlier_frac=0.2)
    (gc, ga), (oc, oa) = _pair(offR, offA, corr, typ, bR=41, bA=41, nGood=40, maxR=0.3, maxA=0.4)
    device.cull("st_cull_st_data", gc)
    oracle32.cull("st_cull_st_data", oc)
    assert gc.nCulled == oc.nCulled
    assert np.array_equal(ga["offR"], oa["offR"]) and np.array_equal(ga["offA"], oa["offA"])


def test_cull_box_above_limit_refused(device):
    offR, offA, corr, typ = make_cull_grid(20, 20, seed=13)
    gc, _ = cull_params(offR, offA, corr, typ, bR=47, bA=9)
    with pytest.raises(RuntimeError, match="box sizes above 45"):
        device.cull("st_cull_st_data", gc)


def test_cull_stats_sigma(device, oracle32):
    offR, offA, corr, typ = make_cull_grid(60, 70, seed=9)
    (gc, ga), (oc, oa) = _pair(offR, offA, corr, typ)
    for _ in range(3):
        oracle32.cull("st_cull_st_data", oc)
    ga["offR"][:] = oa["offR"]
    ga["offA"][:] = oa["offA"]
    device.cull("st_cull_stats", gc)
    oracle32.cull("st_cull_stats", oc)
    assert np.array_equal(ga["sigmaR"] < -1e9, oa["sigmaR"] < -1e9)
    for k in ("sigmaR", "sigmaA"):
        ok, err = _close(ga[k], oa[k])
        assert ok, (k, err)


@pytest.mark.parametrize("s", [(2, 8), (3, 3), (1, 5)])
def test_cull_smooth(device, oracle32, s):
    offR, offA, corr, typ = make_cull_grid(48, 44, seed=12)
    (gc, ga), (oc, oa) = _pair(offR, offA, corr, typ, sR=s[0], sA=s[1])
    oracle32.cull("st_cull_st_data", oc)
    oracle32.cull("st_cull_stats", oc)
    for k in ("offR", "offA", "sigmaR", "sigmaA"):
        ga[k][:] = oa[k]
    device.cull("st_cull_smooth", gc)
    oracle32.cull("st_cull_smooth", oc)
    assert np.array_equal(ga["offRS"] < -1e9, oa["offRS"] < -1e9)
    for k in ("offRS", "offAS", "sigmaR", "sigmaA"):
        ok, err = _close(ga[k], oa[k])
        assert ok, (k, err)


@pytest.mark.parametrize("thresh", [1, 3, 6])
def test_cull_islands_exact(device, oracle32, thresh):
    rng = np.random.default_rng(20 + thresh)
    valid = rng.random((70, 90)) < 0.5
    RS = np.where(valid, 1.0, -2e9).astype(np.float32)
    (gc, ga), (oc, oa) = _pair(RS, RS, np.ones_like(RS), np.ones(RS.shape, np.uint8), islandThresh=thresh)
    for a in (ga, oa):
        a["offRS"][:] = RS
        a["offAS"][:] = RS
    device.cull("st_cull_islands", gc)
    oracle32.cull("st_cull_islands", oc)
    assert np.array_equal(ga["offRS"], oa["offRS"]) and np.array_equal(ga["offAS"], oa["offAS"])


def test_cull_run_full_grid(device, oracle32):
    """The whole cullst sequence on a grid of the full-scene tracking job
    (821 x 571 points)."""
    offR, offA, corr, typ = make_cull_grid(571, 821, seed=30, outlier_frac=0.08)
    (gc, ga), (oc, oa) = _pair(offR, offA, corr, typ, islandThresh=3)
    device.cull("st_cull_run", gc)
    oracle32.cull("st_cull_run", oc)
    for k in ("offR", "offA", "type"):
        assert np.array_equal(ga[k], oa[k]), k
    assert np.array_equal(ga["offRS"] < -1e9, oa["offRS"] < -1e9)
    for k in ("offRS", "offAS", "sigmaR", "sigmaA"):
        ok, err = _close(ga[k], oa[k])
        assert ok, (k, err)
