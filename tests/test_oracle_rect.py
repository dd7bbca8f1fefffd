"""Rectangular windows (wR != wA, wRa != wAa; parseTrack.c:167-171) in the
oracle: the Hanning taper as the reference's storage leaves it
(speckleTrack.c:1775,1830-1852), the padded correlation surface on
2wA x 2wR, and tracker known answers whose limits differ per axis.  The
expected values are worked out by hand from the reference's index
arithmetic, not recomputed by a second copy of the oracle's loops."""
import ctypes as C

import numpy as np
import pytest

from specklesource_amd import abi
from tests.helpers import periodic_pair, small_params


def rect_params(wa, wr, nr, na, step=None, **kw):
    """small_params with a wA x wR complex window and, unless given, the
    same amplitude window."""
    kw.setdefault("wAa", wa)
    kw.setdefault("wRa", wr)
    return small_params(max(wa, wr), nr, na, step=step, wA=wa, wR=wr, **kw)


def _h1(k, n):
    return 0.5 * (1.0 + np.cos(2.0 * np.pi * k / (n - 1)))


def _row_weight(r, wa):
    """computeHanning's wi of row r: rows wA/2 + i and wA/2 - 1 - i get the
    weight of loop index i."""
    i = r - wa // 2 if r >= wa // 2 else wa // 2 - 1 - r
    return _h1(i, wa)


def _hanning(o, wa, wr):
    lib = o.lib
    lib.orc_hanning2.argtypes = [C.c_int, C.c_int, C.c_void_p]
    h = np.full((wa, wr), np.nan, np.float32)
    lib.orc_hanning2(wa, wr, h.ctypes.data)
    return h


# ----------------------------------------------------------- Hanning
def test_hanning_wide_window_spills_into_the_previous_row(oracle32):
    """wA = 32, wR = 64: columns are indexed j1 = 16 + j, j2 = 15 - j for
    j < 32 on rows of 64 floats (speckleTrack.c:1843-1847).  So row r holds
    wi(r) wj(c - 16) at c in [16, 48) and wi(r) wj(15 - c) at c < 16, while
    the negative j2 of row r + 1 land in row r's columns [48, 64) as
    wi(r + 1) wj(79 - c).  Row 31's last 16 cells are never written (0), and
    row 0's negative writes fall before the block (dropped, not wrapped)."""
    wa, wr = 32, 64
    h = _hanning(oracle32, wa, wr)
    want = np.zeros((wa, wr))
    for r in range(wa):
        for c in range(wr):
            if 16 <= c < 48:
                want[r, c] = _row_weight(r, wa) * _h1(c - 16, wr)
            elif c < 16:
                want[r, c] = _row_weight(r, wa) * _h1(15 - c, wr)
            elif r < wa - 1:
                want[r, c] = _row_weight(r + 1, wa) * _h1(79 - c, wr)
    assert np.allclose(h, want.astype(np.float32), rtol=0, atol=1e-7)
    assert np.all(h[wa - 1, 48:] == 0)
    # not the separable taper a column index of wR / 2 would give
    d = lambda n: np.concatenate([np.arange(n // 2)[::-1], np.arange(n // 2)])  # noqa: E731
    assert not np.allclose(h, np.outer(_h1(d(wa), wa), _h1(d(wr), wr)), atol=1e-3)


def test_hanning_tall_window_spills_into_the_next_row(oracle32):
    """wA = 64, wR = 32: j1 = 32 + j lands one row further down at column j
    (value wi(r - 1) wj(c) for c < 16), j2 = 31 - j stays in the row at
    columns [16, 32) (value wi(r) wj(31 - c)).  Row 0's first 16 cells are
    never written and row 63's spill past the block is dropped."""
    wa, wr = 64, 32
    h = _hanning(oracle32, wa, wr)
    want = np.zeros((wa, wr))
    for r in range(wa):
        for c in range(wr):
            if c >= 16:
                want[r, c] = _row_weight(r, wa) * _h1(31 - c, wr)
            elif r >= 1:
                want[r, c] = _row_weight(r - 1, wa) * _h1(c, wr)
    assert np.allclose(h, want.astype(np.float32), rtol=0, atol=1e-7)
    assert np.all(h[0, :16] == 0)


# ------------------------------------------------ correlation surface
@pytest.mark.parametrize("wa,wr", [(16, 32), (32, 16)])
@pytest.mark.parametrize("az", [0, 5])
def test_rectangular_surface_is_band_limited_correlation(oracle64, wa, wr, az):
    """cmpPSWithPad + the 2wA x 2wR inverse (speckleTrack.c:604-643,664):
    C[n, m] = sum over k in [-wA/2, wA/2), l in [-wR/2, wR/2) of
    X[(k + az) mod wA, l mod wR] exp(2 pi i (k n / 2wA + l m / 2wR)) with
    X = F1 conj(F2), fftshifted.  Swapping the axes' sizes or rotating the
    range axis fails it."""
    lib = oracle64.lib
    lib.orc_corr_surface2.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p]
    rng = np.random.default_rng(10 * az + wa)
    p1 = rng.standard_normal((wa, wr)) + 1j * rng.standard_normal((wa, wr))
    p2 = rng.standard_normal((wa, wr)) + 1j * rng.standard_normal((wa, wr))
    mag = np.zeros((2 * wa, 2 * wr))
    lib.orc_corr_surface2(p1.ctypes.data, p2.ctypes.data, wa, wr, az, mag.ctypes.data)
    X = np.fft.fft2(p1) * np.conj(np.fft.fft2(p2))
    ka = np.arange(-wa // 2, wa // 2)
    kr = np.arange(-wr // 2, wr // 2)
    Xr = X[(ka[:, None] + az) % wa, kr[None, :] % wr]
    EA = np.exp(2j * np.pi * np.outer(ka, np.arange(2 * wa)) / (2 * wa))
    ER = np.exp(2j * np.pi * np.outer(kr, np.arange(2 * wr)) / (2 * wr))
    ref = np.abs(np.fft.fftshift(EA.T @ Xr @ ER))
    assert np.allclose(mag, ref, rtol=1e-10, atol=1e-7 * ref.max())
    # the Doppler rotation acts on azimuth only
    if az:
        Xw = X[ka[:, None] % wa, (kr[None, :] + az) % wr]
        assert not np.allclose(mag, np.abs(np.fft.fftshift(EA.T @ Xw @ ER)), rtol=1e-3)


# ------------------------------------------------ tracker known answers
@pytest.mark.parametrize("wa,wr", [(32, 64), (64, 32)])
@pytest.mark.parametrize("shift", [(5, -3), (-6, 4)])
def test_rectangular_complex_match_exact_for_circular_shift(oracle32, oracle64, wa, wr, shift):
    """Windows of an image tiling one wA x wR pattern are circular shifts
    of each other: the exact offset and, without the taper, a normalised
    correlation of 1 (cMax over (wR wA)^2, speckleTrack.c:704-705)."""
    dr, da = shift
    n = 8 * max(wa, wr)
    img1, img2 = periodic_pair((wa, wr), n, n, dr, da, seed=wa + 3)
    par = rect_params(wa, wr, n, n, step=max(wa, wr) // 2, hanningFlag=0, rShiftPoly=[dr, 0, 0],
                      aShiftPoly=[da, 0, 0])
    for o in (oracle32, oracle64):
        r = o.speckle_track(par, img1, img2, diag=True)
        assert (r.type == abi.ST_CMATCH).all(), (o.prec, np.unique(r.type))
        assert np.all(r.offR == dr) and np.all(r.offA == da)
        assert np.allclose(r.corr, 1.0, atol=2e-4)


@pytest.mark.parametrize("wa,wr", [(32, 64), (64, 32)])
def test_rectangular_acceptance_limits_follow_each_axis(oracle32, oracle64, wa, wr):
    """finalizeCmpxMatch's limits are 0.125 wR OS in range and 0.125 wA OS
    in azimuth (speckleTrack.c:731-732): with wA = 32, wR = 64 a 15 px
    range shift is a complex match and an 8 px azimuth shift is not (the
    amplitude try, limit 0.15 wAa 2 = 9.6 px, takes it); the swapped window
    swaps the outcome."""
    n = 8 * max(wa, wr)
    short = min(wa, wr) // 4
    long_ok = max(wa, wr) // 4 - 1
    cases = []
    if wr > wa:
        cases = [((long_ok, 0), abi.ST_CMATCH), ((0, short), abi.ST_AMPMATCH), ((short - 1, 0), abi.ST_CMATCH)]
    else:
        cases = [((0, long_ok), abi.ST_CMATCH), ((short, 0), abi.ST_AMPMATCH), ((0, short - 1), abi.ST_CMATCH)]
    for (dr, da), want in cases:
        img1, img2 = periodic_pair((wa, wr), n, n, dr, da, seed=17)
        par = rect_params(wa, wr, n, n, step=max(wa, wr), hanningFlag=0)
        for o in (oracle32, oracle64):
            r = o.speckle_track(par, img1, img2, diag=True)
            assert np.all(r.type == want), (o.prec, dr, da, np.unique(r.type))
            assert np.all(r.offR == dr) and np.all(r.offA == da)


@pytest.mark.parametrize("wa,wr", [(32, 64), (64, 32)])
def test_rectangular_amplitude_correlation_of_a_circular_shift_is_one(oracle32, oracle64, wa, wr):
    """ampTrackFast's normalisation on a wAa x wRa window: cMax =
    |Q| / ((2 wAa)^2 (2 wRa)^2 64) / sqrt(p1 p2) (speckleTrack.c:591-592)
    is exactly 1 for circularly shifted windows; a square W^4 in place of
    WA^2 WR^2 would give 4 or 1/4."""
    dr, da = 3, -2
    n = 8 * max(wa, wr)
    img1, img2 = periodic_pair((wa, wr), n, n, dr, da, seed=29)
    par = rect_params(wa, wr, n, n, step=max(wa, wr), noComplex=1, maxTries=1, rShiftPoly=[dr, 0, 0],
                      aShiftPoly=[da, 0, 0])
    for o, tol in ((oracle64, 1e-9), (oracle32, 2e-5)):
        r = o.speckle_track(par, img1, img2, diag=True)
        ok = r.type == abi.ST_AMPMATCH
        assert ok.all(), (o.prec, np.unique(r.type))
        assert np.all(r.offR == dr) and np.all(r.offA == da)
        c = r.diag["cmax"][1][ok]
        assert np.abs(c - 1.0).max() < tol, (o.prec, c.min(), c.max())


@pytest.mark.parametrize("wa,wr", [(32, 64), (64, 32)])
def test_rectangular_strackw_integer_shift(oracle32, oracle64, wa, wr):
    """corrTrackFast on a wAa x wRa window of an image tiling one wA x wR
    pattern: every window pair is a circular shift, so the interpolated
    peak sits on the integer shift (within the 1/32 px quantum of its
    refinement) with a correlation near 1."""
    from tests.test_oracle_strackw import stw_params

    dr, da = 3, -2
    n = 8 * max(wa, wr)
    img1, img2 = periodic_pair((wa, wr), n, n, dr, da, seed=31)
    par = stw_params(max(wa, wr), n, n, edge=8, navg=2, step=max(wa, wr), wAa=wa, wRa=wr, wA=wa, wR=wr)
    for o in (oracle32, oracle64):
        g = o.corr_track_fast(par, img1, img2)
        assert (g.type == 2).all(), (o.prec, np.unique(g.type))
        assert np.abs(g.offR - dr).max() <= 1.0 / 32 + 1e-6
        assert np.abs(g.offA - da).max() <= 1.0 / 32 + 1e-6
        assert (g.corr > 0.9).all()
