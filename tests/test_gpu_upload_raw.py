"""Uploads from samples as the reference's files store them
(st_pair_upload_raw): big-endian c64 under MSB, little-endian c64 under
LSB, big-endian ci16 always (speckleTrack.c:1380-1390).  The GPU byte
swap must reproduce the native-order pair exactly."""
import numpy as np
import pytest

from specklesource_amd import abi, configs, track
from tests.helpers import small_params

pytestmark = pytest.mark.gpu


def _pair(seed, nr=333, na=211):
    rng = np.random.default_rng(seed)
    a = (rng.standard_normal((na, nr)) + 1j * rng.standard_normal((na, nr))).astype(np.complex64)
    b = (rng.standard_normal((na, nr)) + 1j * rng.standard_normal((na, nr))).astype(np.complex64)
    return a, b


@pytest.mark.parametrize("order", [1, 0])
def test_c64_raw_upload_equals_native(device, order):
    a, b = _pair(order)  # odd sizes: a tail shorter than 16 bytes
    dt = ">c8" if order == 1 else "<c8"
    pid = device.upload_pair_raw(a.astype(dt).tobytes(), a.shape, b.astype(dt).tobytes(), b.shape,
                                 abi.ST_SAMPLE_C64, order)
    try:
        ga, gb = device.download_pair(pid, 0, a.shape[0], a.shape[1])
    finally:
        device.release_pair(pid)
    assert ga.tobytes() == a.tobytes() and gb.tobytes() == b.tobytes()


def test_ci16_raw_upload_tracks_like_native(device):
    s = configs._spec(400, 400, 21, dr0=1.2, da0=0.4, gmin=0.6, gmax=0.6)
    a, b = track.synth_host(s)
    a16 = np.zeros(a.shape + (2,), np.int16)
    b16 = np.zeros_like(a16)
    a16[..., 0], a16[..., 1] = np.round(a.real * 300), np.round(a.imag * 300)
    b16[..., 0], b16[..., 1] = np.round(b.real * 300), np.round(b.imag * 300)
    p = small_params(32, 400, 400, step=16, rShiftPoly=[1, 0, 0])
    pid = device.upload_pair(a16, b16, abi.ST_SAMPLE_CI16)
    try:
        g0 = device.speckle_track(pid, p)
    finally:
        device.release_pair(pid)
    # int16 SLCs are big-endian in the reference's files whatever the byte-order flag says
    for order in (0, 1):
        pid = device.upload_pair_raw(a16.astype(">i2").tobytes(), a.shape, b16.astype(">i2").tobytes(), b.shape,
                                     abi.ST_SAMPLE_CI16, order)
        try:
            g1 = device.speckle_track(pid, p)
        finally:
            device.release_pair(pid)
        for f in ("offR", "offA", "corr", "type"):
            assert np.array_equal(getattr(g0, f), getattr(g1, f)), f


def test_raw_upload_rejects_bad_order(device):
    a, b = _pair(3, 16, 16)
    with pytest.raises(RuntimeError):
        device.upload_pair_raw(a.tobytes(), a.shape, b.tobytes(), b.shape, abi.ST_SAMPLE_C64, 5)


def test_swap_throughput_full_scene_pair(device):
    """The byte swap of a 20000 x 14000 c64 pair (4.5 GB) runs at HBM
    rates; prints GB/s from the library's event timing."""
    nr, na = 20000, 14000
    raw = np.zeros(nr * na * 8, np.uint8)
    device.set_profiling(True)
    try:
        pid = device.upload_pair_raw(raw, (na, nr), raw, (na, nr), abi.ST_SAMPLE_C64, 1)
        st = device.last_stats()
        names = {st.kernelNames[i].value.decode(): st.kernelMs[i] for i in range(len(st.kernelNames))}
        device.release_pair(pid)
    finally:
        device.set_profiling(False)
    ms = names["k_bswap"]
    gbs = 2 * 2 * raw.nbytes / (ms * 1e-3) / 1e9  # read + write of both images
    print("\nbyte swap of a %.1f GB pair: %.2f ms, %.0f GB/s" % (2 * raw.nbytes / 1e9, ms, gbs))
    assert ms > 0
