"""The closed-form cases that pin the oracle, run through the HIP library:
amplitude correlation of circularly shifted windows is 1, the complex
acceptance limit is a quarter window (strict), the small amplitude try's
limit is 0.15 * wa * 2, on every path that takes those windows."""
import numpy as np
import pytest

from specklesource_amd import abi
from tests.helpers import periodic_pair, small_params

pytestmark = pytest.mark.gpu


def _track(device, par, a, b):
    pid = device.upload_pair(a, b)
    try:
        return device.speckle_track(pid, par, diag=True)
    finally:
        device.release_pair(pid)


@pytest.mark.parametrize("gain", [1.0, 4.0])
def test_amplitude_correlation_of_a_circular_shift_is_one(device, gain):
    w, dr, da = 32, 2, -3
    img1, img2 = periodic_pair(w, 8 * w, 8 * w, dr, da, seed=7)
    img2 = (img2 * gain).astype(np.complex64)
    par = small_params(w, 8 * w, 8 * w, step=w, noComplex=1, maxTries=1, rShiftPoly=[dr, 0, 0],
                       aShiftPoly=[da, 0, 0])
    g = _track(device, par, img1, img2)
    assert np.all(g.type == abi.ST_AMPMATCH)
    assert np.all(g.offR == dr) and np.all(g.offA == da)
    assert np.abs(g.diag["cmax"][1] - 1.0).max() < 2e-5


@pytest.mark.parametrize("w", [16, 32, 64, 128])
@pytest.mark.parametrize("axis", ["r", "a"])
def test_complex_acceptance_limit_is_a_quarter_window(device, w, axis):
    """w = 16, 32, 64 run the fused kernel, w = 128 the staged path."""
    for s, want in ((w // 4 - 1, abi.ST_CMATCH), (w // 4, abi.ST_AMPMATCH)):
        for sign in (1, -1):
            dr, da = (sign * s, 0) if axis == "r" else (0, sign * s)
            img1, img2 = periodic_pair(w, 6 * w, 6 * w, dr, da, seed=11)
            par = small_params(w, 6 * w, 6 * w, step=w, hanningFlag=0)
            g = _track(device, par, img1, img2)
            assert np.all(g.type == want), (w, s, sign, np.unique(g.type))
            assert np.all(g.offR == dr) and np.all(g.offA == da)


@pytest.mark.parametrize("w", [16, 32])
def test_complex_acceptance_limit_on_the_matrix_core_path(device, w):
    for s, want in ((w // 4 - 1, abi.ST_CMATCH), (w // 4, abi.ST_AMPMATCH)):
        img1, img2 = periodic_pair(w, 8 * w, 8 * w, s, -s, seed=12)
        par = small_params(w, 8 * w, 8 * w, step=w, hanningFlag=0, precision=abi.ST_PREC_MFMA)
        g = _track(device, par, img1, img2)
        assert np.all(g.type == want), (w, s, np.unique(g.type))
        assert np.all(g.offR == s) and np.all(g.offA == -s)


@pytest.mark.parametrize("axis", ["r", "a"])
def test_amplitude_acceptance_limit(device, axis):
    w = 32
    for s, want in ((9, abi.ST_AMPMATCH), (10, abi.ST_BAD), (-9, abi.ST_AMPMATCH), (-10, abi.ST_BAD)):
        dr, da = (s, 0) if axis == "r" else (0, s)
        img1, img2 = periodic_pair(w, 8 * w, 8 * w, dr, da, seed=13)
        par = small_params(w, 8 * w, 8 * w, step=w, noComplex=1, maxTries=1)
        g = _track(device, par, img1, img2)
        assert np.all(g.type == want), (s, np.unique(g.type))
        if want == abi.ST_AMPMATCH:
            assert np.all(g.offR == dr) and np.all(g.offA == da)
