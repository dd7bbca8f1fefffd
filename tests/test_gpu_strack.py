"""GPU parity of st_speckle_track (strack) against the oracle.

Integer-peak locations and match types must agree with the FP32-faithful
oracle except where the oracle's own diagnostics show a near-tie (top two
correlation values within TIE) or a value within THRESH_TOL of an
acceptance threshold; those are counted.  Sub-pixel offsets are compared
with the FP64 oracle: RMS <= 0.02 px per axis.  Correlation relative
error against FP64 is reported and bounded by 1e-3 (median)."""
import numpy as np
import pytest

from specklesource_amd import abi, configs, track
from tests.helpers import compare_strack, periodic_pair, small_params

pytestmark = pytest.mark.gpu

RMS_TOL = 0.02


def _run_all(device, oracle32, oracle64, par, a, b, stype=abi.ST_SAMPLE_C64):
    pid = device.upload_pair(a, b, stype)
    try:
        g = device.speckle_track(pid, par, diag=True)
    finally:
        device.release_pair(pid)
    o32 = oracle32.speckle_track(par, a, b, stype, diag=True)
    o64 = oracle64.speckle_track(par, a, b, stype, diag=True)
    rep, probs = compare_strack(par, g, o32, o64)
    print("\n", rep)
    return g, o32, o64, rep, probs


def _check(rep, probs, rms=True):
    assert not probs, probs[:10]
    assert rep["peak_agree"] >= 0.999 or rep["attempts"] - rep["near_ties"] == round(rep["peak_agree"] * rep["attempts"])
    if rms and "rms_r_vs_f64" in rep:
        assert rep["rms_r_vs_f64"] <= RMS_TOL and rep["rms_a_vs_f64"] <= RMS_TOL
        assert rep["corr_relerr_median"] < 1e-3


@pytest.mark.parametrize("w", [16, 32, 64])
def test_circular_shift_known_answer(device, w):
    img1, img2 = periodic_pair(w, 8 * w, 8 * w, 3, -2, seed=w)
    par = small_params(w, 8 * w, 8 * w, step=w // 2, hanningFlag=0, rShiftPoly=[3, 0, 0], aShiftPoly=[-2, 0, 0])
    pid = device.upload_pair(img1, img2)
    r = device.speckle_track(pid, par)
    device.release_pair(pid)
    assert (r.type == abi.ST_CMATCH).all()
    assert np.all(r.offR == 3) and np.all(r.offA == -2)
    assert np.allclose(r.corr, 1.0, atol=2e-4)


def test_parity_full_scene_shape(device, oracle32, oracle64):
    """cfg2-shaped scene (w=64 complex + amplitude retries, ridge flow,
    coherence 0.3-0.9, zero-filled block), scaled to 1000 x 700."""
    s, p = configs.cfg2(scale=20)
    a, b = track.synth_host(s)
    g, o32, o64, rep, probs = _run_all(device, oracle32, oracle64, p, a, b)
    assert (g.type == 0).sum() > 0 and (g.type == 1).sum() > 0
    _check(rep, probs)


def test_parity_amplitude_only(device, oracle32, oracle64):
    """cfg1-shaped (-noComplex, w=32, large retry 64) on 512 x 1024."""
    s, p = configs.cfg1(scale=4)
    a, b = track.synth_host(s)
    g, o32, o64, rep, probs = _run_all(device, oracle32, oracle64, p, a, b)
    assert (g.type == abi.ST_AMPMATCH).mean() > 0.5
    _check(rep, probs)


@pytest.mark.parametrize("w", [24, 48])
def test_parity_amplitude_radix3_windows(device, oracle32, oracle64, w):
    """-noComplex with windows of 24 and 48 (transforms of 48 / 96 / 192
    points, radix 3 stages) and the doubled-window retries."""
    s = configs._spec(480, 480, 200 + w, dr0=0.6, drR=1e-3, da0=-1.1, daA=2e-3, gmin=0.3, gmax=0.85, gLr=300,
                      gLa=250)
    p = small_params(w, 480, 480, step=w // 2, noComplex=1, rShiftPoly=[0.6, 1e-3, 0], aShiftPoly=[-1.1, 0, 2e-3])
    a, b = track.synth_host(s)
    g, o32, o64, rep, probs = _run_all(device, oracle32, oracle64, p, a, b)
    assert g.nAmp > 0.3 * g.type.size
    _check(rep, probs)


@pytest.mark.parametrize("w", [16, 32])
def test_parity_window_sizes(device, oracle32, oracle64, w):
    s = configs._spec(480, 480, 100 + w, dr0=0.8, drR=1e-3, da0=-1.3, daA=2e-3, gmin=0.35, gmax=0.85,
                      gLr=300, gLa=250)
    p = small_params(w, 480, 480, step=w // 2, rShiftPoly=[0.8, 1e-3, 0], aShiftPoly=[-1.3, 0, 2e-3])
    a, b = track.synth_host(s)
    g, o32, o64, rep, probs = _run_all(device, oracle32, oracle64, p, a, b)
    _check(rep, probs)


def test_parity_complex_int16(device, oracle32, oracle64):
    s = configs._spec(400, 400, 21, dr0=1.2, da0=0.4, gmin=0.6, gmax=0.6)
    a, b = track.synth_host(s)
    a16 = np.zeros(a.shape + (2,), np.int16)
    b16 = np.zeros_like(a16)
    a16[..., 0] = np.round(a.real * 300)
    a16[..., 1] = np.round(a.imag * 300)
    b16[..., 0] = np.round(b.real * 300)
    b16[..., 1] = np.round(b.imag * 300)
    p = small_params(32, 400, 400, step=16, rShiftPoly=[1, 0, 0])
    g, o32, o64, rep, probs = _run_all(device, oracle32, oracle64, p, a16, b16, abi.ST_SAMPLE_CI16)
    _check(rep, probs)


@pytest.mark.parametrize("w", [16, 64, 128])
def test_parity_complex_int16_fused_windows(device, oracle32, oracle64, w):
    """Integer (CI16) samples through the fused kernels of the other window
    classes: their 16-bit loads are instantiations of their own."""
    n = 480 if w <= 64 else 1100
    s = configs._spec(n, n, 40 + w, dr0=1.1, drR=1e-3, da0=-0.6, daA=1e-3, gmin=0.5, gmax=0.85, gLr=300,
                      gLa=250)
    a, b = track.synth_host(s)
    a16 = np.zeros(a.shape + (2,), np.int16)
    b16 = np.zeros_like(a16)
    a16[..., 0] = np.round(a.real * 300)
    a16[..., 1] = np.round(a.imag * 300)
    b16[..., 0] = np.round(b.real * 300)
    b16[..., 1] = np.round(b.imag * 300)
    p = small_params(w, n, n, step=w // 2 if w <= 64 else 48, rShiftPoly=[1.1, 1e-3, 0], aShiftPoly=[-0.6, 0, 1e-3])
    device.set_profiling(True)
    try:
        g, o32, o64, rep, probs = _run_all(device, oracle32, oracle64, p, a16, b16, abi.ST_SAMPLE_CI16)
        names = _kernel_names(device)
    finally:
        device.set_profiling(False)
    assert names.get("k_cpx_fused", 0) > 0
    assert (g.type == abi.ST_CMATCH).mean() > 0.5
    _check(rep, probs)


def test_parity_initial_grid_with_holes(device, oracle32, oracle64):
    """findImage2Pos from an initial-offset grid; grid holes keep the
    previous point's window position (speckleTrack.c:823-826)."""
    s = configs._spec(400, 400, 22, dr0=2.4, da0=-1.7, gmin=0.7, gmax=0.7)
    a, b = track.synth_host(s)
    p = small_params(32, 400, 400, step=16, polyShift=0)
    gr = np.full((10, 10), 2.4, np.float32)
    ga = np.full((10, 10), -1.7, np.float32)
    gr[3:5, 2:6] = -2e9
    ga[3:5, 2:6] = -2e9
    p.initDr = abi.ptr(gr, abi.C.c_float)
    p.initDa = abi.ptr(ga, abi.C.c_float)
    p.initNr = p.initNa = 10
    p.initR0, p.initA0, p.initDeltaR, p.initDeltaA = 0.0, 0.0, 40.0, 40.0
    g, o32, o64, rep, probs = _run_all(device, oracle32, oracle64, p, a, b)
    assert np.array_equal(g.diag["r2"], o32.diag["r2"]) and np.array_equal(g.diag["a2"], o32.diag["a2"])
    _check(rep, probs)


def test_parity_mask_and_flags(device, oracle32, oracle64):
    s = configs._spec(400, 400, 23, dr0=0.7, da0=1.1, gmin=0.3, gmax=0.7, gLr=200, gLa=300)
    a, b = track.synth_host(s)
    mask = np.ones((20, 20), np.uint8)
    mask[5:9, :] = 0
    for kw in (dict(maxTries=1), dict(hanningFlag=0), dict(checkAzFocus=1, azDefocusThresh=1.2),
               dict(scaleFactor=2, rStart=40, aStart=40, deltaR=8, deltaA=8)):
        p = small_params(32, 400, 400, step=16, rShiftPoly=[1, 0, 0], aShiftPoly=[1, 0, 0])
        p.mask = abi.ptr(mask, abi.C.c_uint8)
        p.maskNr = p.maskNa = 20
        p.maskNrl = p.maskNal = 20
        for k, v in kw.items():
            setattr(p, k, v)
        g, o32, o64, rep, probs = _run_all(device, oracle32, oracle64, p, a, b)
        _check(rep, probs)


def test_parity_zero_filled_block(device, oracle32, oracle64):
    s = configs._spec(400, 400, 24, dr0=0.5, da0=0.5, gmin=0.7, gmax=0.7, maskR0=100, maskR1=220,
                      maskA0=80, maskA1=300)
    a, b = track.synth_host(s)
    p = small_params(32, 400, 400, step=12)
    g, o32, o64, rep, probs = _run_all(device, oracle32, oracle64, p, a, b)
    assert np.array_equal(g.type == 0, o32.type == 0)
    _check(rep, probs)


def test_parity_baseline_phase(device, oracle32, oracle64):
    """phaseCorrectBaseline (speckleTrack.c:868-907) with a non-zero
    baseline and per-range look-angle tables: patch 1 is multiplied by
    exp(i phi) before the complex match."""
    s = configs._spec(400, 400, 25, dr0=0.9, da0=-0.6, gmin=0.7, gmax=0.8)
    a, b = track.synth_host(s)
    p = small_params(32, 400, 400, step=16, rShiftPoly=[1, 0, 0], aShiftPoly=[-1, 0, 0])
    j = np.arange(400, dtype=np.float64)
    theta = 0.6 + 4e-5 * j
    sinT = np.sin(theta)
    cosT = np.cos(theta)
    overR = 1.0 / (8.0e5 + 2.3 * j)
    p.baseFlag = 1
    p.Bn, p.Bp, p.dBn, p.dBp, p.dBnQ, p.dBpQ = 50.0, 20.0, 0.02, -0.01, 0.001, 0.0005
    p.lambda_ = 0.0555
    p.nSlpA1 = 400
    p.sinThetaD = abi.ptr(sinT, abi.C.c_double)
    p.cosThetaD = abi.ptr(cosT, abi.C.c_double)
    p.overRange = abi.ptr(overR, abi.C.c_double)
    g, o32, o64, rep, probs = _run_all(device, oracle32, oracle64, p, a, b)
    # the phase ramp costs about half of the complex matches: both paths run
    assert g.nComplex > 50 and g.nAmp > 50
    _check(rep, probs)


def _kernel_names(device):
    st = device.last_stats()
    return {st.kernelNames[i].value.decode(): st.kernelLaunches[i] for i in range(len(st.kernelNames))}


@pytest.mark.parametrize("w", [16, 32, 64, 128])
def test_staged_path_parity(device, oracle32, oracle64, monkeypatch, w):
    """The staged complex path (range transforms, then k_cpx_colx,
    k_cpx_rowmag and k_peak_merge) forced at the fused kernels' window
    sizes, against the same oracle bars."""
    monkeypatch.setenv("SPECKLE_AMD_PATH", "staged")
    n = 480 if w <= 64 else 1100
    s = configs._spec(n, n, 300 + w, dr0=0.8, drR=1e-3, da0=-1.3, daA=2e-3, gmin=0.35, gmax=0.85,
                      gLr=300, gLa=250)
    p = small_params(w, n, n, step=w // 2, rShiftPoly=[0.8, 1e-3, 0], aShiftPoly=[-1.3, 0, 2e-3])
    a, b = track.synth_host(s)
    device.set_profiling(True)
    try:
        g, o32, o64, rep, probs = _run_all(device, oracle32, oracle64, p, a, b)
        names = _kernel_names(device)
    finally:
        device.set_profiling(False)
    assert names.get("peak_refine", 0) > 0 and names.get("k_cpx_fused", 0) == 0
    _check(rep, probs)


def test_circular_shift_known_answer_w128(device):
    """w = 128 (the lane-quad fused kernel): circular shifts give exact
    integer offsets and correlation 1."""
    w = 128
    img1, img2 = periodic_pair(w, 6 * w, 6 * w, 5, -3, seed=w)
    par = small_params(w, 6 * w, 6 * w, step=w // 2, hanningFlag=0, rShiftPoly=[5, 0, 0], aShiftPoly=[-3, 0, 0])
    pid = device.upload_pair(img1, img2)
    r = device.speckle_track(pid, par)
    device.release_pair(pid)
    assert (r.type == abi.ST_CMATCH).all()
    assert np.all(r.offR == 5) and np.all(r.offA == -3)
    assert np.allclose(r.corr, 1.0, atol=2e-4)


def test_parity_window_128(device, oracle32, oracle64):
    """w = 128 with a sheared offset field and partial coherence: the
    largest window of the batched uneven-pairs workload."""
    s = configs._spec(1100, 1100, 128, dr0=2.3, drR=2e-3, da0=-1.7, daA=1e-3, gmin=0.4, gmax=0.9, gLr=400,
                      gLa=300)
    p = small_params(128, 1100, 1100, step=48, rShiftPoly=[2.3, 2e-3, 0], aShiftPoly=[-1.7, 0, 1e-3])
    a, b = track.synth_host(s)
    device.set_profiling(True)
    try:
        g, o32, o64, rep, probs = _run_all(device, oracle32, oracle64, p, a, b)
        names = _kernel_names(device)
    finally:
        device.set_profiling(False)
    assert names.get("k_cpx_fused", 0) > 0 and names.get("peak_refine", 0) == 0
    assert (g.type == abi.ST_CMATCH).mean() > 0.5
    _check(rep, probs)


def test_parity_decorrelated_amplitude_retries(device, oracle32, oracle64):
    """Low coherence at w = 64: most complex matches fail, so the amplitude
    tries at W = 128 and the doubled-window tries at W = 256 carry the job
    (the separable interpolation, packed forward transform, conjugate
    separation and paired-row inverse of the amplitude path)."""
    s = configs._spec(720, 720, 91, dr0=0.7, drR=1e-3, da0=-0.4, daA=1e-3, gmin=0.12, gmax=0.2)
    p = small_params(64, 720, 720, step=32, rShiftPoly=[0.7, 1e-3, 0], aShiftPoly=[-0.4, 0, 1e-3])
    a, b = track.synth_host(s)
    g, o32, o64, rep, probs = _run_all(device, oracle32, oracle64, p, a, b)
    assert g.nAmp > 30 and g.nAmpL > 60
    _check(rep, probs)
