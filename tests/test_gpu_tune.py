"""st_autotune: the tuner writes its cache in the documented text format,
the choices it keeps come back through st_tune_get and st_tune_load, and a
call with ST_PREC_AUTO gives exactly what the chosen precision gives."""
import numpy as np
import pytest

from specklesource_amd import abi, configs, track
from tests.helpers import small_params

pytestmark = pytest.mark.gpu


def test_autotune_cache_and_auto_precision(device, tmp_path):
    cache = tmp_path / "tune.txt"
    device.autotune([16, 32], str(cache))
    lines = [ln.split() for ln in cache.read_text().splitlines() if ln and not ln.startswith("#")]
    assert [int(ln[1]) for ln in lines] == [16, 32]
    for ln in lines:
        assert ln[0].startswith("gfx") and ln[2] in ("fp32", "mfma") and int(ln[3]) >= 0 and int(ln[4]) >= 0
        assert float(ln[5]) > 0
    got = {w: device.tune(w) for w in (16, 32)}
    assert all(v is not None for v in got.values())
    device.tune_load(str(cache))
    assert {w: device.tune(w)[:3] for w in (16, 32)} == {w: v[:3] for w, v in got.items()}
    for w in (16, 32):
        s = configs._spec(320, 320, 800 + w, dr0=0.7, da0=-0.4, gmin=0.5, gmax=0.9)
        pid = device.synth_pair(s)
        try:
            auto = small_params(w, 320, 320, step=w // 2, rShiftPoly=[0.7, 0, 0], aShiftPoly=[-0.4, 0, 0],
                                precision=abi.ST_PREC_AUTO)
            chosen = small_params(w, 320, 320, step=w // 2, rShiftPoly=[0.7, 0, 0], aShiftPoly=[-0.4, 0, 0],
                                  precision=got[w][0])
            a = device.speckle_track(pid, auto)
            b = device.speckle_track(pid, chosen)
            for f in ("offR", "offA", "corr", "type"):
                assert np.array_equal(getattr(a, f), getattr(b, f)), (w, f)
        finally:
            device.release_pair(pid)
