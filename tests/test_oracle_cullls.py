"""Pin the Cullls oracle (orc_cull_ls_*: Cullls/cullLSData.c, cullLSStats.c,
cullLSSmooth.c, cullLSIslands.c) against a literal sequential pass, hand-made
known answers and closed forms.

The slow-velocity bit, the no-data-to-use bit, NODATA and the Landsat NOVER
live in headers outside the reference tree; the tests use the values below
and check the behaviour each one controls."""
import numpy as np
import pytest

from specklesource_amd import track
from tests.test_oracle_cull import _seq_cull_bruteforce

SLOW, NODATA_BIT = 0x10, 0x20
NODATA = np.float32(-2e9)
LSKW = dict(slow_vel_mask=SLOW, no_data_to_use_mask=NODATA_BIT, no_data=-2.0e9, nover=10, deltaT=16.0, dx=15.0,
            dy=15.0)


def _ls(X, Y, typ, **kw):
    k = dict(LSKW)
    k.update(kw)
    return track.cull_ls_params(X, Y, typ, **k)


def _ls_grid(ny, nx, seed, outlier_frac=0.1, slow_region=True):
    rng = np.random.default_rng(seed)
    i = np.arange(ny)[:, None]
    j = np.arange(nx)[None, :]
    X = (0.3 + 0.01 * i - 0.02 * j + 0.05 * rng.standard_normal((ny, nx))).astype(np.float32)
    Y = (-0.2 + 0.015 * j + 0.05 * rng.standard_normal((ny, nx))).astype(np.float32)
    out = rng.random((ny, nx)) < outlier_frac
    X[out] += rng.choice([-1, 1], out.sum()) * rng.uniform(0.25, 3.0, out.sum()).astype(np.float32)
    hole = rng.random((ny, nx)) < 0.05
    X[hole] = NODATA
    Y[hole] = NODATA
    typ = np.ones((ny, nx), np.uint8)
    if slow_region:
        typ[: ny // 2] |= SLOW
    typ[hole] |= NODATA_BIT
    return X, Y, typ


def _seq_ls(X, Y, typ, b=9, nGood=17, maxX=0.75, maxY=1.0, deltaT=16.0, dx=15.0, dy=15.0, noData=-2e9):
    """Literal sequential cullLSData pass in Python."""
    X, Y = X.copy(), Y.copy()
    ny, nx = X.shape
    for i in range(ny):
        for j in range(nx):
            box = (slice(max(0, i - b // 2), min(ny, i + b // 2 + 1)), slice(max(0, j - b // 2), min(nx, j + b // 2 + 1)))
            v = (X[box] > 1 - 2e9) & (Y[box] > 1 - 2e9)
            n = int(v.sum())
            slow = bool(np.all(typ[box][v] & SLOW)) if n else True
            if n > nGood and float(Y[i, j]) > noData + 1:
                k = max(n // 2, 1)
                mx = np.sort(X[box][v])[k - 1]
                my = np.sort(Y[box][v])[k - 1]
                if slow:
                    lx = max(40.0 / 365.0 * deltaT / dx, 0.2)
                    ly = max(40.0 / 365.0 * deltaT / dy, 0.2)
                else:
                    lx, ly = float(np.float32(maxX)), float(np.float32(maxY))
                if abs(float(np.float32(Y[i, j] - my))) > ly or abs(float(np.float32(X[i, j] - mx))) > lx:
                    X[i, j] = Y[i, j] = NODATA
            else:
                X[i, j] = Y[i, j] = NODATA
    return X, Y


@pytest.mark.parametrize("shape,seed", [((30, 40), 1), ((45, 23), 2)])
def test_ls_pass_matches_literal_sequential_pass(oracle32, shape, seed):
    X, Y, typ = _ls_grid(*shape, seed=seed)
    cp, arrs = _ls(X, Y, typ)
    Xs, Ys = X, Y
    for _ in range(3):
        oracle32.cull("st_cull_ls_data", cp)
        Xs, Ys = _seq_ls(Xs, Ys, typ)
        assert np.array_equal(arrs["X"], Xs) and np.array_equal(arrs["Y"], Ys)
    # the first pass recorded the attempt statistics
    assert cp.nAttempt == int(((typ & NODATA_BIT) == 0).sum())
    assert cp.nMatch == int(((X > 1 - 2e9) & (Y > 1 - 2e9)).sum())


def test_slow_limit_known_answer(oracle32):
    """A flat field with one X outlier of +0.5: with every box point slow,
    the limit is max(40/365 * 16 / 15, 0.2) = 0.2 and it is culled; one
    non-slow point in its box switches to maxX = 0.75 and it stays; an
    outlier of +1.0 goes either way."""
    ny, nx = 15, 15
    X = np.zeros((ny, nx), np.float32)
    Y = np.zeros((ny, nx), np.float32)
    X[7, 7] = 0.5
    slow = np.full((ny, nx), 1 | SLOW, np.uint8)
    cp, arrs = _ls(X, Y, slow)
    oracle32.cull("st_cull_ls_data", cp)
    assert arrs["X"][7, 7] == NODATA and cp.nCulled == 1
    mixed = slow.copy()
    mixed[5, 5] = 1
    cp, arrs = _ls(X, Y, mixed)
    oracle32.cull("st_cull_ls_data", cp)
    assert arrs["X"][7, 7] == np.float32(0.5) and cp.nCulled == 0
    X[7, 7] = 1.0
    cp, arrs = _ls(X, Y, mixed)
    oracle32.cull("st_cull_ls_data", cp)
    assert arrs["X"][7, 7] == NODATA
    # a long interval raises the slow limit: 40/365 * 100 / 15 = 0.73
    X[7, 7] = 0.5
    cp, arrs = _ls(X, Y, slow, deltaT=100.0)
    oracle32.cull("st_cull_ls_data", cp)
    assert arrs["X"][7, 7] == np.float32(0.5)


def test_nodata_cell_is_culled(oracle32):
    """A cell whose Y is not above NODATA + 1 is culled although X and Y are
    valid offsets and agree with the median: a constant field Y = -1000 is
    culled entirely with NODATA = -1000 and kept entirely with -1001.5."""
    ny, nx = 12, 12
    X = np.zeros((ny, nx), np.float32)
    Y = np.full((ny, nx), -1000.0, np.float32)
    typ = np.ones((ny, nx), np.uint8)
    cp, arrs = _ls(X, Y, typ, no_data=-1000.0)
    oracle32.cull("st_cull_ls_data", cp)
    assert cp.nCulled == ny * nx and (arrs["X"] == NODATA).all()
    cp, arrs = _ls(X, Y, typ, no_data=-1001.5)
    oracle32.cull("st_cull_ls_data", cp)
    assert cp.nCulled == 0 and (arrs["Y"] == np.float32(-1000.0)).all()


def test_ls_stats_plane_closed_form(oracle64):
    """An exact plane leaves zero residuals: sigma = sqrt(1e-5 + (1/NOVER)^2 / 12)
    wherever 9 or more good points surround a good cell."""
    ny, nx = 20, 25
    i = np.arange(ny, dtype=np.float32)[:, None]
    j = np.arange(nx, dtype=np.float32)[None, :]
    X = (0.5 + 0.03 * i - 0.01 * j).astype(np.float32)
    Y = np.broadcast_to(-1.0 + 0.02 * j, (ny, nx)).astype(np.float32)
    cp, arrs = _ls(X, Y, np.ones((ny, nx), np.uint8), nover=10)
    oracle64.cull("st_cull_ls_stats", cp)
    want = np.float32(np.sqrt(np.float32(np.float32(1e-5) + 0.01 / 12)))
    assert np.allclose(arrs["sigmaX"], want, rtol=2e-6) and np.allclose(arrs["sigmaY"], want, rtol=2e-6)
    cp, arrs = _ls(X, Y, np.ones((ny, nx), np.uint8), nover=20)
    oracle64.cull("st_cull_ls_stats", cp)
    want = np.float32(np.sqrt(np.float32(np.float32(1e-5) + 0.0025 / 12)))
    assert np.allclose(arrs["sigmaX"], want, rtol=2e-6)


def test_ls_smooth_closed_form(oracle64):
    """Unweighted box means (3 x 3 interior, clipped at the edges) and sigma
    divided by sqrt(count)."""
    rng = np.random.default_rng(4)
    ny, nx = 10, 12
    X = rng.standard_normal((ny, nx)).astype(np.float32)
    Y = rng.standard_normal((ny, nx)).astype(np.float32)
    cp, arrs = _ls(X, Y, np.ones((ny, nx), np.uint8))
    arrs["sigmaX"][:] = 2.0
    arrs["sigmaY"][:] = 3.0
    oracle64.cull("st_cull_ls_smooth", cp)
    for (a, b) in [(0, 0), (4, 5), (9, 11), (0, 6)]:
        box = (slice(max(0, a - 1), min(ny, a + 2)), slice(max(0, b - 1), min(nx, b + 2)))
        n = X[box].size
        assert arrs["XS"][a, b] == pytest.approx(X[box].astype(np.float64).mean(), rel=1e-6)
        assert arrs["YS"][a, b] == pytest.approx(Y[box].astype(np.float64).mean(), rel=1e-6)
        assert arrs["sigmaX"][a, b] == pytest.approx(2.0 / np.sqrt(n), rel=1e-6)
        assert arrs["sigmaY"][a, b] == pytest.approx(3.0 / np.sqrt(n), rel=1e-6)


def test_ls_islands_equal_cullst_islands(oracle32):
    from tests.helpers import cull_params
    rng = np.random.default_rng(9)
    XS = np.full((30, 30), NODATA, np.float32)
    for _ in range(12):
        a, b, h, w = rng.integers(0, 26), rng.integers(0, 26), rng.integers(1, 6), rng.integers(1, 6)
        XS[a:a + h, b:b + w] = 1.0
    YS = XS.copy()
    cp, arrs = _ls(np.zeros_like(XS), np.zeros_like(XS), np.ones(XS.shape, np.uint8), islandThresh=3)
    arrs["XS"][:] = XS
    arrs["YS"][:] = YS
    oracle32.cull("st_cull_ls_islands", cp)
    sp, sarr = cull_params(np.zeros_like(XS), np.zeros_like(XS), np.ones_like(XS), np.ones(XS.shape, np.uint8),
                           islandThresh=3)
    sarr["offRS"][:] = XS
    sarr["offAS"][:] = YS
    oracle32.cull("st_cull_islands", sp)
    assert np.array_equal(arrs["XS"], sarr["offRS"]) and np.array_equal(arrs["YS"], sarr["offAS"])
    assert (arrs["XS"] != XS).any()


def test_ls_data_without_slow_points_equals_cullst_pass(oracle32):
    """No slow bits and no NODATA cells: cullLSData is cullSTData's median
    pass with (maxX, maxY) in place of (maxR, maxA)."""
    X, Y, typ = _ls_grid(30, 35, seed=5, slow_region=False)
    ok = (X > 1 - 2e9)
    X = np.where(ok, X, np.float32(0)).astype(np.float32)
    Y = np.where(ok, Y, np.float32(0)).astype(np.float32)
    cp, arrs = _ls(X, Y, np.ones(X.shape, np.uint8))
    oracle32.cull("st_cull_ls_data", cp)
    R, A = _seq_cull_bruteforce(X, Y, maxA=1.0, maxR=0.75)
    assert np.array_equal(arrs["X"], R) and np.array_equal(arrs["Y"], A)
