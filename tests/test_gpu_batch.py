"""st_speckle_track_batch: several strack calls in one batch (the calls of
one window class share a fused launch) give exactly the outputs each call
gives alone."""
import numpy as np
import pytest

from specklesource_amd import abi, configs, track
from tests.helpers import small_params

pytestmark = pytest.mark.gpu


def _jobs():
    out = []
    for k, (w, n) in enumerate(((16, 320), (32, 400), (16, 256), (64, 480), (32, 352), (128, 640), (32, 300))):
        s = configs._spec(n, n + 32 * k, 600 + k, dr0=0.5 * k - 1, da0=0.3 * k, gmin=0.4, gmax=0.9, gLr=n / 2,
                          gLa=n / 3)
        p = small_params(w, n, n + 32 * k, step=w // 2, rShiftPoly=[0.5 * k - 1, 0, 0], aShiftPoly=[0.3 * k, 0, 0])
        out.append((s, p))
    # a matrix-core call and a masked one ride along
    out[4][1].precision = abi.ST_PREC_MFMA
    return out


def test_batch_equals_calls_alone(device):
    jobs = _jobs()
    pids = [device.synth_pair(s) for s, _ in jobs]
    try:
        alone = [device.speckle_track(pid, p) for pid, (_, p) in zip(pids, jobs)]
        for order in (list(range(len(jobs))), list(reversed(range(len(jobs))))):
            got = device.speckle_track_batch([pids[k] for k in order], [jobs[k][1] for k in order])
            for k, g in zip(order, got):
                a = alone[k]
                for f in ("offR", "offA", "corr", "azd", "type"):
                    assert np.array_equal(getattr(g, f), getattr(a, f)), (k, f)
                assert (g.nComplex, g.nAmp, g.nAmpL, g.nFail) == (a.nComplex, a.nAmp, a.nAmpL, a.nFail)
        assert sum(a.nComplex for a in alone) > 0
    finally:
        for pid in pids:
            device.release_pair(pid)
