"""Rectangular windows on the GPU (wR != wA and wRa != wAa,
parseTrack.c:167-171): strack on the staged complex and amplitude paths
with the reference's Hanning storage (speckleTrack.c:1775,1830-1852), and
strackw, against the oracle on the same inputs at the parity bars of the
square-window tests."""
import numpy as np
import pytest

from specklesource_amd import abi, configs, track
from tests.helpers import periodic_pair
from tests.test_gpu_strack import _check, _run_all
from tests.test_oracle_rect import rect_params

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("wa,wr", [(32, 64), (64, 32)])
def test_rectangular_circular_shift_known_answer(device, wa, wr):
    n = 8 * max(wa, wr)
    img1, img2 = periodic_pair((wa, wr), n, n, 5, -3, seed=wa + 3)
    par = rect_params(wa, wr, n, n, step=max(wa, wr) // 2, hanningFlag=0, rShiftPoly=[5, 0, 0],
                      aShiftPoly=[-3, 0, 0])
    pid = device.upload_pair(img1, img2)
    try:
        r = device.speckle_track(pid, par)
    finally:
        device.release_pair(pid)
    assert (r.type == abi.ST_CMATCH).all()
    assert np.all(r.offR == 5) and np.all(r.offA == -3)
    assert np.allclose(r.corr, 1.0, atol=2e-4)


def _scene(wa, wr, seed, n=600, gmin=0.3, gmax=0.9):
    s = configs._spec(n, n, seed, dr0=0.9, drR=1e-3, da0=-1.2, daA=2e-3, gmin=gmin, gmax=gmax, gLr=n / 2,
                      gLa=n / 3)
    return s, track.synth_host(s)


@pytest.mark.parametrize("wa,wr", [(32, 64), (64, 32)])
@pytest.mark.parametrize("hanning", [0, 1])
def test_parity_rectangular_windows(device, oracle32, oracle64, wa, wr, hanning):
    """Complex matches with amplitude retries (small and doubled windows)
    on a sheared field with coherence 0.05-0.5; with the taper on, patch 2
    is weighted by the reference's mis-indexed Hanning block."""
    s, (a, b) = _scene(wa, wr, 400 + wa + hanning, gmin=0.05, gmax=0.5)
    p = rect_params(wa, wr, 600, 600, step=24, hanningFlag=hanning, rShiftPoly=[0.9, 1e-3, 0],
                    aShiftPoly=[-1.2, 0, 2e-3])
    g, o32, o64, rep, probs = _run_all(device, oracle32, oracle64, p, a, b)
    assert (g.type == abi.ST_CMATCH).mean() > 0.4 and g.nAmp + g.nAmpL > 0
    _check(rep, probs)


@pytest.mark.parametrize("wa,wr", [(32, 64), (64, 32)])
def test_parity_rectangular_amplitude_windows(device, oracle32, oracle64, wa, wr):
    """-noComplex with a wAa x wRa amplitude window (and the doubled window
    on the retries), the small try estimating its own Doppler shift."""
    s, (a, b) = _scene(wa, wr, 450 + wa, n=640)
    p = rect_params(32, 32, 640, 640, step=24, wAa=wa, wRa=wr, noComplex=1, rShiftPoly=[0.9, 1e-3, 0],
                    aShiftPoly=[-1.2, 0, 2e-3])
    g, o32, o64, rep, probs = _run_all(device, oracle32, oracle64, p, a, b)
    assert g.nAmp > 0.3 * g.type.size
    _check(rep, probs)


def test_parity_rectangular_gauss(device, oracle32, oracle64):
    """-gauss on a 32 x 64 window: the 2-D fit on the 64 x 128 surface."""
    s, (a, b) = _scene(32, 64, 470)
    p = rect_params(32, 64, 600, 600, step=32, gaussFlag=1, rShiftPoly=[0.9, 1e-3, 0],
                    aShiftPoly=[-1.2, 0, 2e-3])
    g, o32, o64, rep, probs = _run_all(device, oracle32, oracle64, p, a, b)
    assert (g.type == abi.ST_CMATCH).mean() > 0.3
    _check(rep, probs, rms=False)


def test_rectangular_batch_equals_calls_alone(device):
    """A batch mixing a rectangular call (staged path) with square fused
    calls gives each call's results alone."""
    s, (a, b) = _scene(32, 64, 480)
    jobs = [rect_params(32, 64, 600, 600, step=32, rShiftPoly=[0.9, 0, 0]),
            rect_params(32, 32, 600, 600, step=16, rShiftPoly=[0.9, 0, 0]),
            rect_params(64, 32, 600, 600, step=32, rShiftPoly=[0.9, 0, 0])]
    pid = device.upload_pair(a, b)
    try:
        alone = [device.speckle_track(pid, p) for p in jobs]
        got = device.speckle_track_batch([pid] * len(jobs), jobs)
    finally:
        device.release_pair(pid)
    for k, (x, y) in enumerate(zip(got, alone)):
        for f in ("offR", "offA", "corr", "type"):
            assert np.array_equal(getattr(x, f), getattr(y, f)), (k, f)
    assert all(r.nComplex > 0 for r in alone)


@pytest.mark.parametrize("wa,wr", [(64, 128), (128, 64)])
def test_parity_rectangular_strackw(device, oracle32, oracle64, wa, wr):
    """strackw on a wAa x wRa window (getCorrPatchesFast forces wR = wRa,
    wA = wAa; corrTrackFast.c:322-323), edge pad 8, 2 x 2 boxcar."""
    from tests.test_gpu_strackw import _check as stw_check, _run as stw_run, _scene as stw_scene, compare

    s, par = stw_scene(520, 520, 60 + wa, max(wa, wr), 8, 2, step=48, wAa=wa, wRa=wr, wA=wa, wR=wr)
    a, b = track.synth_host(s)
    g = stw_run(device, par, a, b)
    o32 = oracle32.corr_track_fast(par, a, b, diag=True)
    o64 = oracle64.corr_track_fast(par, a, b, diag=True)
    rep, probs = compare(g, o32, o64)
    print("\n", rep)
    assert g.nAmp > 0.5 * g.type.size
    stw_check(rep, probs)
