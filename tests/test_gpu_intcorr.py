"""GPU parity of the interferogram phase correction (phaseCorrectInt,
speckleTrack.c:909-1000) against the oracle.

Image 1 carries an interferogram fringe; the correction multiplies patch 1
by conj(y) / |y| before the complex match, so the matches that the fringe
removes come back.  The same bars as the other strack parity tests apply
(integer peaks and match types against the FP32-faithful oracle up to
near-ties, sub-pixel RMS against FP64).  The cases cover the fused kernel
at w = 16, 32, 64, patches above 64 samples, the staged path, the correction after the baseline
ramp, spectra with two kept bins and flat (chirp) spectra whose kept bins
exceed the per-point record, and the sizes the GPU path refuses."""
import ctypes as C

import numpy as np
import pytest

from specklesource_amd import abi, track
from tests.helpers import compare_strack, fringe_scene
from tests.test_gpu_strack import _check, _kernel_names, _run_all

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("w,nrl,nal,kfa,kfr", [(16, 1, 2, 1, 3), (32, 2, 8, 2, 3), (64, 2, 4, 3, 7)])
def test_int_parity_fused(device, oracle32, oracle64, w, nrl, nal, kfa, kfr):
    par, a, b, intf, ps = fringe_scene(24 * w if w < 64 else 12 * w, w, nrl, nal, kfa, kfr, seed=w)
    device.set_profiling(True)
    try:
        g, o32, o64, rep, probs = _run_all(device, oracle32, oracle64, par, a, b)
        names = _kernel_names(device)
    finally:
        device.set_profiling(False)
    n = par.nR * par.nA
    assert names.get("k_cpx_fused", 0) > 0 and names.get("k_int_bins", 0) == 1
    assert g.nComplex > 0.7 * n, (g.nComplex, n)
    _check(rep, probs)


@pytest.mark.parametrize("w,kfa,kfr", [(96, 2, 5), (128, 3, 7)])
def test_int_parity_large_patch(device, oracle32, oracle64, w, kfa, kfr):
    """Correction patches above 64 samples a side (w = 96 and 128 without
    looks, ps = w), whose spectra go through a global workspace."""
    par, a, b, intf, ps = fringe_scene(12 * w, w, 1, 1, kfa, kfr, seed=w + 1)
    assert ps == w
    device.set_profiling(True)
    try:
        g, o32, o64, rep, probs = _run_all(device, oracle32, oracle64, par, a, b)
        names = _kernel_names(device)
    finally:
        device.set_profiling(False)
    assert names.get("k_int_bins", 0) >= 1
    assert g.nComplex > 0.7 * par.nR * par.nA, (g.nComplex, par.nR * par.nA)
    _check(rep, probs)


def test_int_off_loses_matches(device):
    """Known answer of the scene: without the interferogram the fringe
    (3 cycles over the window) leaves few complex matches."""
    par, a, b, intf, ps = fringe_scene(640, 32, 2, 8, 2, 3, seed=3)
    pid = device.upload_pair(a, b)
    try:
        on = device.speckle_track(pid, par)
        par.intf = None
        off = device.speckle_track(pid, par)
    finally:
        device.release_pair(pid)
    n = par.nR * par.nA
    assert on.nComplex > 0.7 * n and off.nComplex < 0.3 * n, (on.nComplex, off.nComplex, n)


def test_int_parity_staged(device, oracle32, oracle64, monkeypatch):
    monkeypatch.setenv("SPECKLE_AMD_PATH", "staged")
    par, a, b, intf, ps = fringe_scene(640, 32, 2, 8, 2, 3, seed=11)
    device.set_profiling(True)
    try:
        g, o32, o64, rep, probs = _run_all(device, oracle32, oracle64, par, a, b)
        names = _kernel_names(device)
    finally:
        device.set_profiling(False)
    assert names.get("peak_refine", 0) > 0 and names.get("k_cpx_fused", 0) == 0
    assert g.nComplex > 0.7 * par.nR * par.nA
    _check(rep, probs)


def test_int_after_baseline(device, oracle32, oracle64):
    """Baseline ramp, then the interferogram correction, on patch 1 (the
    order of complexMatch, speckleTrack.c:362-363)."""
    n = 768
    par, a, b, intf, ps = fringe_scene(n, 64, 2, 4, 3, 7, seed=21)
    j = np.arange(n, dtype=np.float64)
    theta = 0.6 + 4e-5 * j
    sinT, cosT, overR = np.sin(theta), np.cos(theta), 1.0 / (8.0e5 + 2.3 * j)
    par.baseFlag = 1
    par.Bn, par.Bp, par.dBn, par.dBp, par.dBnQ, par.dBpQ = 5.0, 2.0, 0.02, -0.01, 0.001, 0.0005
    par.lambda_ = 0.0555
    par.nSlpA1 = n
    par.sinThetaD = abi.ptr(sinT, C.c_double)
    par.cosThetaD = abi.ptr(cosT, C.c_double)
    par.overRange = abi.ptr(overR, C.c_double)
    g, o32, o64, rep, probs = _run_all(device, oracle32, oracle64, par, a, b)
    assert g.nComplex > 0
    _check(rep, probs)


@pytest.mark.parametrize("path", ["fused", "staged"])
def test_int_many_kept_bins(device, oracle32, oracle64, monkeypatch, path):
    """Two equal fringes over the right half (two kept bins per point) and a
    quadratic-phase pattern over the lower half, whose flat spectrum keeps
    all 256 bins: those points overflow the per-point record and take the
    second pass.

    The flat-spectrum patches give y samples close to zero, where the
    FP32-faithful oracle's phase is rounding noise (the FP32 and FP64
    oracles differ there at 2 of 1024 samples and pick different first
    peaks on 5 of the failed attempts); the GPU evaluates y in FP64, so its
    integer peaks are compared with the FP64 oracle, its match types with
    both."""
    if path == "staged":
        monkeypatch.setenv("SPECKLE_AMD_PATH", "staged")
    par, a, b, intf, ps = fringe_scene(640, 32, 2, 8, 2, 3, seed=17, noise=0.0, second=(5, 1, 1.0), chirp=True)
    device.set_profiling(True)
    try:
        g, o32, o64, rep, probs = _run_all(device, oracle32, oracle64, par, a, b)
        names = _kernel_names(device)
    finally:
        device.set_profiling(False)
    assert names.get("k_int_bins", 0) == 2
    assert (g.type == o32.type).all() and (g.type == o64.type).all()
    rep64, probs64 = compare_strack(par, g, o64, o64)
    assert not probs64 and rep64["peak_agree"] == 1.0, probs64[:10]
    assert rep64["rms_r_vs_f64"] <= 0.02 and rep64["rms_a_vs_f64"] <= 0.02


def test_int_refused_sizes(device):
    """ps above 128 (w = 192 without looks) is refused as unsupported; a
    window not covered by ps x looks (w = 40 with 3 x 3 looks: ps = 13,
    Pa = 39) as an argument error, as the oracle does."""
    img = np.zeros((600, 600), np.complex64)
    pid = device.upload_pair(img, img)
    try:
        intf = np.ones((300, 600), np.float32)
        from tests.helpers import small_params
        p = small_params(192, 600, 600)
        p.intf = abi.ptr(intf, C.c_float)
        p.intNa, p.intNr, p.intNrl, p.intNal = 300, 300, 1, 1
        with pytest.raises(RuntimeError, match=r"\(-3\).*interferogram patch above 128"):
            device.speckle_track(pid, p)
        p = small_params(40, 600, 600)
        p.intf = abi.ptr(intf, C.c_float)
        p.intNa, p.intNr, p.intNrl, p.intNal = 300, 300, 3, 3
        with pytest.raises(RuntimeError, match=r"\(-1\)"):
            device.speckle_track(pid, p)
    finally:
        device.release_pair(pid)
