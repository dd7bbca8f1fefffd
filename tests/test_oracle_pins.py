"""Pins of the oracle rules that the tracker-level known answers leave
open: the amplitude correlation's normalisation, the acceptance limits of
both match kinds, findCPeak's scan order among equal maxima and getShifts'
rule for holes in the initial-offset grid.  Each expected value is worked
out by hand from the reference's rule (cited per test), not recomputed by
a second copy of the oracle's code."""
import ctypes as C

import numpy as np
import pytest

from specklesource_amd import abi, track
from tests.helpers import periodic_pair, small_params


# ----------------------------------------- amplitude normalisation
@pytest.mark.parametrize("shift", [(2, -3), (-4, 1)])
@pytest.mark.parametrize("gain", [1.0, 4.0])
def test_amplitude_correlation_of_a_circular_shift_is_one(oracle32, oracle64, shift, gain):
    """ampTrackFast normalises the oversampled peak as
    cMax = |Q| / ((2 wAa)^2 (2 wRa)^2 64) / sqrt(p1 p2)
    (speckleTrack.c:591-592).  For windows that are circular shifts of
    each other the interpolated, detected, mean-removed amplitudes v1, v2
    are circular shifts too, so the cross-correlation at the peak lag is
    N sum v1^2 with N = (2 wa)^2 from the unnormalised forward and inverse
    transforms (N^2 p), the 8 x 8 -> 96 x 96 interpolation reproduces it
    x 64 at the sample, and p1 = p2 = p: cMax = 1 exactly.  Dropping the
    x 64, squaring N once or taking p as a sum instead of a mean moves it
    by orders of magnitude; scaling image 2 must not move it at all."""
    w = 32
    dr, da = shift
    img1, img2 = periodic_pair(w, 8 * w, 8 * w, dr, da, seed=7)
    img2 = (img2 * gain).astype(np.complex64)
    par = small_params(w, 8 * w, 8 * w, step=w, noComplex=1, maxTries=1, rShiftPoly=[dr, 0, 0],
                       aShiftPoly=[da, 0, 0])
    for o, tol in ((oracle64, 1e-9), (oracle32, 2e-5)):
        r = o.speckle_track(par, img1, img2, diag=True)
        ok = r.type == abi.ST_AMPMATCH
        assert ok.all()
        assert np.all(r.offR == dr) and np.all(r.offA == da)
        c = r.diag["cmax"][1][ok]
        assert np.abs(c - 1.0).max() < tol, (o.prec, c.min(), c.max())


# ---------------------------------------------------- acceptance limits
@pytest.mark.parametrize("axis", ["r", "a"])
@pytest.mark.parametrize("sign", [1, -1])
def test_complex_acceptance_limit_is_a_quarter_window(oracle32, oracle64, axis, sign):
    """finalizeCmpxMatch accepts |shift| < 0.125 * w * OS = w / 4 px, strict
    (speckleTrack.c:725-733).  With no initial offset, windows that are
    circular shifts by w/4 - 1 px match as complex; at exactly w/4 px the
    complex match is refused and the amplitude try (limit 0.15 wa 2 =
    9.6 px at wa = 32) takes the point."""
    w = 32
    for s, want in ((w // 4 - 1, abi.ST_CMATCH), (w // 4, abi.ST_AMPMATCH)):
        dr, da = (sign * s, 0) if axis == "r" else (0, sign * s)
        img1, img2 = periodic_pair(w, 8 * w, 8 * w, dr, da, seed=11)
        par = small_params(w, 8 * w, 8 * w, step=w, hanningFlag=0)
        for o in (oracle32, oracle64):
            r = o.speckle_track(par, img1, img2, diag=True)
            assert np.all(r.type == want), (o.prec, s, np.unique(r.type))
            assert np.all(r.offR == dr) and np.all(r.offA == da)


@pytest.mark.parametrize("axis", ["r", "a"])
def test_amplitude_acceptance_limit(oracle32, oracle64, axis):
    """The small amplitude try is accepted when |shift| < 0.15 * wa * 2
    (speckleTrack.c:255,289-297): 9 px passes, 10 px fails at wa = 32
    (one try allowed, so a refused point stays type 0)."""
    w = 32
    for s, want in ((9, abi.ST_AMPMATCH), (10, abi.ST_BAD), (-9, abi.ST_AMPMATCH), (-10, abi.ST_BAD)):
        dr, da = (s, 0) if axis == "r" else (0, s)
        img1, img2 = periodic_pair(w, 8 * w, 8 * w, dr, da, seed=13)
        par = small_params(w, 8 * w, 8 * w, step=w, noComplex=1, maxTries=1)
        for o in (oracle32, oracle64):
            r = o.speckle_track(par, img1, img2, diag=True)
            assert np.all(r.type == want), (o.prec, s, np.unique(r.type))
            if want == abi.ST_AMPMATCH:
                assert np.all(r.offR == dr) and np.all(r.offA == da)
            # the try ran and found the shifted peak either way
            assert np.all(r.diag["peakI"][1] > -1000)


# ------------------------------------------------- findCPeak scan order
def _find_cpeak(o, mag_shifted):
    """Run the oracle's findCPeak on the surface whose fftshifted
    magnitude is mag_shifted (W_A x W_R)."""
    WA, WR = mag_shifted.shape
    Cs = np.roll(np.roll(mag_shifted, WA // 2, axis=0), WR // 2, axis=1)  # undo the fftshift
    Cz = np.ascontiguousarray(Cs.astype(o.dtype))
    real = np.float32 if o.prec == "f32" else np.float64
    mag = np.zeros((WA, WR), real)
    i, j, c, snr = C.c_int(), C.c_int(), C.c_double(), C.c_double()
    o.lib.orc_find_cpeak.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.POINTER(C.c_int),
                                     C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    o.lib.orc_find_cpeak(Cz.ctypes.data, WR, WA, mag.ctypes.data, C.byref(i), C.byref(j), C.byref(c),
                         C.byref(snr))
    assert np.allclose(mag, mag_shifted)
    _find_cpeak.snr = snr.value
    return i.value, j.value, c.value


# Equal maxima at fftshifted positions (row, column) of a 32 x 32 surface
# (H = 16) and the winner.  The reference scans base cells (i, j) in [0,
# H)^2 row-major and, per base cell, (i, j), (i+H, j+H), (i+H, j),
# (i, j+H), keeping the first strict maximum (speckleTrack.c:750-789).
# A plain row-major scan of the shifted surface picks the other answer in
# every case below.
TIE_CASES = [
    # base (1, 7) of (17, 23) comes before base (2, 5)
    ([(5, 6), (17, 23)], (17, 23)),
    # one base cell, four quadrants: (i+H, j+H) before (i+H, j) and (i, j+H)
    ([(19, 20), (19, 4), (3, 20)], (19, 20)),
    # (i+H, j) before (i, j+H) although row 3 comes first row-major
    ([(3, 20), (19, 4)], (19, 4)),
    # same row of the shifted surface: base column 2 (col 18) before base column 9
    ([(6, 9), (6, 18)], (6, 18)),
    # the border rule still applies to the winner: row 2 < NFAST / 2
    ([(2, 10), (20, 10)], (-1000, -1000)),
]


@pytest.mark.parametrize("case", range(len(TIE_CASES)))
def test_find_cpeak_resolves_equal_maxima_in_the_reference_scan_order(oracle32, oracle64, case):
    pos, want = TIE_CASES[case]
    rng = np.random.default_rng(case)
    mag = rng.random((32, 32)) * 0.5
    for p in pos:
        mag[p] = 2.0
    for o in (oracle32, oracle64):
        i, j, c = _find_cpeak(o, mag)
        assert (i, j) == want, (o.prec, pos, (i, j))
        assert c == (-1.0 if want[0] == -1000 else 2.0)


# --------------------------------------------- getShifts' hole rule
def _grid_params(dr, da):
    nA, nR = dr.shape
    par = track.default_params(rStart=0, aStart=0, deltaR=1, deltaA=1, nR=1, nA=1, wR=16, wA=16, wRa=16,
                               wAa=16, polyShift=0)
    keep = (np.ascontiguousarray(dr, np.float32), np.ascontiguousarray(da, np.float32))
    par.initDr = abi.ptr(keep[0], C.c_float)
    par.initDa = abi.ptr(keep[1], C.c_float)
    par.initNr, par.initNa = nR, nA
    par.initR0, par.initA0, par.initDeltaR, par.initDeltaA = 100.0, 50.0, 10.0, 20.0
    return par, keep


def _shift(o, par, r, a):
    o.lib.orc_get_shifts.argtypes = [C.POINTER(abi.TrackParams), C.c_double, C.c_double,
                                     C.POINTER(C.c_double), C.POINTER(C.c_double)]
    dr, da = C.c_double(), C.c_double()
    o.lib.orc_get_shifts(C.byref(par), r, a, C.byref(dr), C.byref(da))
    return dr.value, da.value


HOLE = -2.0e9


def test_get_shifts_bilinear_inside_and_nearest_on_the_last_row_and_column(oracle32, oracle64):
    """speckleTrack.c:1870-1914: grid coordinates (r - r0) / dR, (a - a0) / dA;
    outside the grid -LARGEINT; on the last row or column the node value
    itself; elsewhere bilinear in the cell (hand-computed values)."""
    dr = np.array([[1, 2, 3], [4, 5, 6], [7, 8, 9]], np.float32)
    da = 10 * dr
    par, keep = _grid_params(dr, da)
    for o in (oracle32, oracle64):
        # cell (0, 0) centre: (1 + 2 + 5 + 4) / 4 = 3
        assert _shift(o, par, 105.0, 60.0) == pytest.approx((3.0, 30.0))
        # t = 0.25, u = 0.5 in cell (1, 1): rows 5 6 / 8 9 -> 5.25 + 1.5
        assert _shift(o, par, 112.5, 80.0) == pytest.approx((6.75, 67.5))
        # last column (j = 2): node value, no interpolation
        assert _shift(o, par, 125.0, 55.0) == (3.0, 30.0)
        # last row (i = 2)
        assert _shift(o, par, 111.0, 95.0) == (8.0, 80.0)
        # before the grid origin, and past its end
        assert _shift(o, par, 99.0, 60.0) == (-2e9, -2e9)
        assert _shift(o, par, 131.0, 60.0) == (-2e9, -2e9)


def test_get_shifts_takes_the_largest_corner_next_to_a_hole(oracle32, oracle64):
    """speckleTrack.c:1898-1930: when any of the four corners is at or below
    -0.9 LARGEINT the result is the largest corner (no interpolation); when
    all four are holes both offsets become -LARGEINT, even if the other
    component interpolated fine."""
    dr = np.array([[1, HOLE, 3], [4, 5, 6], [7, 8, 9]], np.float32)
    da = np.array([[10, 20, 30], [40, 50, 60], [70, 80, 90]], np.float32)
    par, keep = _grid_params(dr, da)
    for o in (oracle32, oracle64):
        # cell (0, 0): corners 1, hole, 5, 4 -> 5; da interpolates to 30
        assert _shift(o, par, 105.0, 60.0) == pytest.approx((5.0, 30.0))
        # cell (0, 1): corners hole, 3, 6, 5 -> 6; da (20 + 30 + 60 + 50) / 4
        assert _shift(o, par, 115.0, 60.0) == pytest.approx((6.0, 40.0))
        # cell (1, 0) has no hole: bilinear
        assert _shift(o, par, 105.0, 80.0) == pytest.approx((6.0, 60.0))
    dr2 = np.array([[HOLE, HOLE, 3], [HOLE, HOLE, 6], [7, 8, 9]], np.float32)
    da2 = np.array([[10, 20, 30], [40, 50, 60], [70, 80, 90]], np.float32)
    par2, keep2 = _grid_params(dr2, da2)
    for o in (oracle32, oracle64):
        assert _shift(o, par2, 105.0, 60.0) == (-2e9, -2e9)
    # a hole in da only: dr interpolates, da takes the largest corner
    da3 = np.array([[HOLE, 20, 30], [40, 50, 60], [70, 80, 90]], np.float32)
    par3, keep3 = _grid_params(dr2 * 0 + 1, da3)
    for o in (oracle32, oracle64):
        assert _shift(o, par3, 105.0, 60.0) == pytest.approx((1.0, 50.0))
    # all four da corners are holes: dr is overwritten too
    da4 = np.array([[HOLE, HOLE, 30], [HOLE, HOLE, 60], [70, 80, 90]], np.float32)
    par4, keep4 = _grid_params(dr2 * 0 + 1, da4)
    for o in (oracle32, oracle64):
        assert _shift(o, par4, 105.0, 60.0) == (-2e9, -2e9)


# ------------------------------------------------------- SNR diagnostic
def test_snr_is_peak_over_mean_surface_magnitude(oracle32, oracle64):
    """SNR = the coarse peak over the mean |C| of the whole surface (a
    diagnostic the reference does not compute): a flat surface of ones with
    one peak of 9 on a 32 x 32 surface gives 9 * 1024 / (1023 + 9); a flat
    surface gives 1."""
    mag = np.ones((32, 32))
    mag[20, 12] = 9.0
    for o in (oracle32, oracle64):
        _find_cpeak(o, mag)
        assert _find_cpeak.snr == pytest.approx(9 * 1024 / 1032, rel=1e-6)
        _find_cpeak(o, np.ones((32, 32)) * 3.0)
        assert _find_cpeak.snr == pytest.approx(1.0, rel=1e-6)
