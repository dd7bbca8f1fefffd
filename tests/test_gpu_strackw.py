"""GPU parity of st_corr_track_fast (strackw) against the oracle.

The coarse NCC peak must agree with the FP32-faithful oracle except at a
near-tie (runner-up within TIE of the best, by either side's own
diagnostics); the oversampled peak must agree except where the fine
surface has a near-tie (fineGap < 1e-3).  Offsets are compared with the
FP64 oracle (RMS <= 0.02 px per axis) and are identical to the FP32
oracle's wherever both peaks agree.  Match types are compared exactly."""
import numpy as np
import pytest

from specklesource_amd import abi, configs, track
from tests.helpers import TIE, periodic_pair
from tests.test_oracle_strackw import stw_params

pytestmark = pytest.mark.gpu

RMS_TOL = 0.02
FINE_TIE = 1e-3


def _run(device, par, a, b, stype=abi.ST_SAMPLE_C64):
    pid = device.upload_pair(a, b, stype)
    try:
        return device.corr_track_fast(pid, par, diag=True)
    finally:
        device.release_pair(pid)


def compare(g, o32, o64):
    rep, probs = {}, []
    np.testing.assert_array_equal(g.type, o32.type)
    ran = g.diag["peakI"][0] > -2000
    d32, dg = o32.diag, g.diag
    coarse_same = ran & (d32["peakI"][0] == dg["peakI"][0]) & (d32["peakJ"][0] == dg["peakJ"][0])
    coarse_tie = ran & ~coarse_same & ((d32["peakGap"][0] < TIE) | (dg["peakGap"][0] < TIE))
    for idx in zip(*np.nonzero(ran & ~coarse_same & ~coarse_tie)):
        probs.append(("coarse", idx, int(d32["peakI"][0][idx]), int(d32["peakJ"][0][idx]),
                      int(dg["peakI"][0][idx]), int(dg["peakJ"][0][idx])))
    fine_same = coarse_same & (d32["fineI"][0] == dg["fineI"][0]) & (d32["fineJ"][0] == dg["fineJ"][0])
    fine_tie = coarse_same & ~fine_same & ((d32["fineGap"][0] < FINE_TIE) | (dg["fineGap"][0] < FINE_TIE))
    for idx in zip(*np.nonzero(coarse_same & ~fine_same & ~fine_tie)):
        probs.append(("fine", idx, int(d32["fineI"][0][idx]), int(d32["fineJ"][0][idx]),
                      int(dg["fineI"][0][idx]), int(dg["fineJ"][0][idx]), float(d32["fineGap"][0][idx])))
    rep["patches"] = int(ran.sum())
    rep["coarse_agree"] = float(coarse_same.sum() / max(ran.sum(), 1))
    rep["fine_agree"] = float(fine_same.sum() / max(ran.sum(), 1))
    rep["coarse_ties"] = int(coarse_tie.sum())
    rep["fine_ties"] = int(fine_tie.sum())
    if fine_same.any():
        assert np.array_equal(g.offR[fine_same], o32.offR[fine_same])
        assert np.array_equal(g.offA[fine_same], o32.offA[fine_same])
    ok = (g.offR > -1e9) & (o64.offR > -1e9)
    if ok.any():
        rep["rms_r_vs_f64"] = float(np.sqrt(np.mean((g.offR[ok] - o64.offR[ok]).astype(np.float64) ** 2)))
        rep["rms_a_vs_f64"] = float(np.sqrt(np.mean((g.offA[ok] - o64.offA[ok]).astype(np.float64) ** 2)))
        cr = np.abs(g.corr[ok].astype(np.float64) - o64.corr[ok]) / np.maximum(np.abs(o64.corr[ok]), 1e-12)
        rep["corr_relerr_max"] = float(cr.max())
    rep["counts"] = (g.nAmp, g.nFail, o32.nAmp, o32.nFail)
    return rep, probs


def _check(rep, probs):
    assert not probs, probs[:10]
    if "rms_r_vs_f64" in rep:
        assert rep["rms_r_vs_f64"] <= RMS_TOL and rep["rms_a_vs_f64"] <= RMS_TOL
        assert rep["corr_relerr_max"] < 1e-3


@pytest.mark.parametrize("dr,da", [(3, -2), (-5, 4)])
def test_integer_shift_known_answer(device, dr, da):
    w = 64
    img1, img2 = periodic_pair(w, 400, 400, dr, da, seed=3)
    par = stw_params(w, 400, 400, edge=10, navg=2, step=48)
    g = _run(device, par, img1, img2)
    assert (g.type == 2).all() and g.nAmp == par.nR * par.nA
    q = 1.0 / 32
    assert np.abs(g.offR - dr).max() <= q + 1e-6
    assert np.abs(g.offA - da).max() <= q + 1e-6
    assert (g.corr > 0.9).all()


def _scene(nr, na, seed, w, edge, navg, step, **kw):
    s = configs._spec(nr, na, seed, dr0=1.7, drR=2e-3, da0=-2.4, daA=3e-3, gmin=0.7, gmax=0.9,
                      gLr=nr / 1.2, gLa=na / 1.4)
    par = stw_params(w, nr, na, edge, navg=navg, step=step, rShiftPoly=[1.7, 2e-3, 0],
                     aShiftPoly=[-2.4, 0, 3e-3], **kw)
    return s, par


@pytest.mark.parametrize("w,edge,navg", [(64, 8, 2), (64, 8, 1), (96, 12, 2), (128, 20, 2)])
def test_parity_synthetic(device, oracle32, oracle64, w, edge, navg):
    s, par = _scene(480, 480, 40 + w + navg, w, edge, navg, step=w // 2)
    a, b = track.synth_host(s)
    g = _run(device, par, a, b)
    o32 = oracle32.corr_track_fast(par, a, b, diag=True)
    o64 = oracle64.corr_track_fast(par, a, b, diag=True)
    rep, probs = compare(g, o32, o64)
    print("\n", rep)
    _check(rep, probs)


@pytest.mark.parametrize("navg", [12, 24])
def test_parity_wide_boxcar(device, oracle32, oracle64, navg):
    """Boxcar half-widths up to 24 (k_w_detect's haloed tiles are sized
    per call) on a coarse grid, so the oracle's (2 navg + 1)^2 taps per
    pixel stay quick.  A 49 x 49 boxcar over a 64 px window leaves a flat
    NCC surface on which FP32 itself (the FP32 oracle) lands up to ~0.1 px
    from FP64, so the GPU's offsets are held to the FP32 oracle's (equal
    peaks give equal offsets; differing peaks must be near-ties) and its
    2-D offset error against FP64 to within RMS_TOL of the FP32 oracle's
    (a near-tie resolved the other way moves a point along either axis, so
    the axes are not bounded one by one)."""
    s, par = _scene(480, 480, 70 + navg, 64, 8, navg, step=96)
    a, b = track.synth_host(s)
    g = _run(device, par, a, b)
    o32 = oracle32.corr_track_fast(par, a, b, diag=True)
    o64 = oracle64.corr_track_fast(par, a, b, diag=True)
    rep, probs = compare(g, o32, o64)
    ok = (o32.offR > -1e9) & (o64.offR > -1e9)
    base = [float(np.sqrt(np.mean((x[ok] - y[ok]).astype(np.float64) ** 2)))
            for x, y in ((o32.offR, o64.offR), (o32.offA, o64.offA))]
    print("\n", rep, "fp32 oracle rms vs fp64:", base)
    assert not probs, probs[:10]
    assert np.hypot(rep["rms_r_vs_f64"], rep["rms_a_vs_f64"]) <= np.hypot(*base) + RMS_TOL


def test_boxcar_above_limit_refused(device):
    s, par = _scene(480, 480, 99, 64, 8, 25, step=96)
    a, b = track.synth_host(s)
    with pytest.raises(RuntimeError, match="boxcar half-widths above 24"):
        _run(device, par, a, b)


def test_parity_cfg4_window(device, oracle32, oracle64):
    """The chip-128 strackw window of the sheared-flow workload (wRa = 384,
    edge pad 123 -> 768^2 transforms, 513^2 lags, 5 x 5 boxcar) on a
    1024 x 1024 scene with offsets of several pixels and coherence
    0.3-0.5, 3 x 3 grid points."""
    s = configs._spec(1024, 1024, 5, dr0=3.2, drR=2e-3, da0=-7.6, daA=-1e-3, gmin=0.3, gmax=0.5,
                      gLr=900, gLa=800)
    p = track.default_params(rStart=448, aStart=448, deltaR=64, deltaA=64, nR=3, nA=3, wR=384, wA=384,
                             wRa=384, wAa=384, edgePadR=123, edgePadA=123, navgR=2, navgA=2, noComplex=1,
                             rShiftPoly=[0, 0, 0], aShiftPoly=[0, 0, 0])
    a, b = track.synth_host(s)
    g = _run(device, p, a, b)
    o32 = oracle32.corr_track_fast(p, a, b, diag=True)
    o64 = oracle64.corr_track_fast(p, a, b, diag=True)
    rep, probs = compare(g, o32, o64)
    print("\n", rep)
    _check(rep, probs)
    tr = np.array([[track.synth_offset(s, 448 + 64 * j, 448 + 64 * i) for j in range(3)] for i in range(3)])
    assert np.abs(g.offR - tr[..., 0]).max() < 0.35 and np.abs(g.offA - tr[..., 1]).max() < 0.35


def test_parity_cfg4_chip256_window(device, oracle32, oracle64):
    """The chip-256 strackw window of the sheared-flow workload: wRa = 512
    (chip 256 + 2 x (123 + 5)), 1024^2 transforms, 513^2 lags, 5 x 5
    boxcar (corrTrackFast.c:116-124), on a 1536 x 1536 scene with offsets
    of tens of pixels found from an initial polynomial a few pixels off,
    coherence 0.2-0.4, 2 x 2 grid points."""
    s = configs._spec(1536, 1536, 6, dr0=20.2, drR=2e-3, da0=-35.6, daA=-1e-3, gmin=0.2, gmax=0.4,
                      gLr=1200, gLa=1000)
    p = track.default_params(rStart=640, aStart=640, deltaR=160, deltaA=160, nR=2, nA=2, wR=512, wA=512,
                             wRa=512, wAa=512, edgePadR=123, edgePadA=123, navgR=2, navgA=2, noComplex=1,
                             rShiftPoly=[18.0, 0, 0], aShiftPoly=[-40.0, 0, 0])
    a, b = track.synth_host(s)
    g = _run(device, p, a, b)
    o32 = oracle32.corr_track_fast(p, a, b, diag=True)
    o64 = oracle64.corr_track_fast(p, a, b, diag=True)
    rep, probs = compare(g, o32, o64)
    print("\n", rep)
    _check(rep, probs)
    assert g.nAmp == 4
    tr = np.array([[track.synth_offset(s, 640 + 160 * j, 640 + 160 * i) for j in range(2)] for i in range(2)])
    assert np.abs(g.offR - tr[..., 0]).max() < 0.35 and np.abs(g.offA - tr[..., 1]).max() < 0.35


def test_ci16_samples(device, oracle32, oracle64):
    s, par = _scene(400, 400, 77, 64, 8, 2, step=40)
    a, b = track.synth_host(s)
    sc = 2000.0
    ai = np.stack([np.round(a.real * sc), np.round(a.imag * sc)], -1).astype(np.int16)
    bi = np.stack([np.round(b.real * sc), np.round(b.imag * sc)], -1).astype(np.int16)
    g = _run(device, par, ai, bi, abi.ST_SAMPLE_CI16)
    o32 = oracle32.corr_track_fast(par, ai, bi, abi.ST_SAMPLE_CI16, diag=True)
    o64 = oracle64.corr_track_fast(par, ai, bi, abi.ST_SAMPLE_CI16, diag=True)
    rep, probs = compare(g, o32, o64)
    _check(rep, probs)


def test_initial_offset_grid_and_mask(device, oracle32, oracle64):
    """polyShift = 0 with an initial-offset grid (bilinear, with a hole that
    falls back to the polynomial) and a mask that zeroes a block."""
    import ctypes as C
    s, par = _scene(480, 480, 91, 64, 8, 2, step=32)
    a, b = track.synth_host(s)
    n0 = 9
    gr = np.zeros((n0, n0), np.float32)
    ga = np.zeros((n0, n0), np.float32)
    for i in range(n0):
        for j in range(n0):
            gr[i, j], ga[i, j] = track.synth_offset(s, 16 + 60 * j + 32, 16 + 60 * i + 32)
    gr[4, 4] = ga[4, 4] = -2e9
    mask = np.ones((60, 60), np.uint8)
    mask[10:25, 30:50] = 0
    par.polyShift = 0
    par.initDr = abi.ptr(gr, C.c_float)
    par.initDa = abi.ptr(ga, C.c_float)
    par.initNr = par.initNa = n0
    par.initR0 = par.initA0 = 16
    par.initDeltaR = par.initDeltaA = 60
    par.mask = abi.ptr(mask, C.c_uint8)
    par.maskNr = par.maskNa = 60
    par.maskNrl = par.maskNal = 8
    g = _run(device, par, a, b)
    o32 = oracle32.corr_track_fast(par, a, b, diag=True)
    o64 = oracle64.corr_track_fast(par, a, b, diag=True)
    assert (g.type == 0).any()
    np.testing.assert_array_equal(g.diag["r2"], o32.diag["r2"])
    np.testing.assert_array_equal(g.diag["a2"], o32.diag["a2"])
    rep, probs = compare(g, o32, o64)
    _check(rep, probs)


def test_featureless_data_fails_with_type2(device):
    w = 48
    img = np.ones((240, 240), np.complex64)
    par = stw_params(w, 240, 240, edge=6, step=40)
    g = _run(device, par, img, img)
    assert (g.type == 2).all() and g.nAmp == 0 and g.nFail == par.nR * par.nA
    assert (g.offR == np.float32(-2e9)).all()
