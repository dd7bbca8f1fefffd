"""The serving layer (st_server): jobs of different kinds and sizes run
concurrently give exactly the results they give alone, admission follows
the HBM budget, and job states and timings are reported."""
import numpy as np
import pytest

from specklesource_amd import abi, configs, track
from tests.helpers import small_params
from tests.test_oracle_strackw import stw_params

pytestmark = pytest.mark.gpu


def _jobs():
    """(kind, params, img1, img2, cull kwargs or None) of a mixed workload."""
    out = []
    for k, (w, n, seed) in enumerate(((16, 320, 1), (32, 480, 2), (64, 640, 3))):
        s = configs._spec(n, n, 300 + seed, dr0=0.6 * k, da0=-0.4, gmin=0.5, gmax=0.9, gLr=n / 2, gLa=n / 3)
        a, b = track.synth_host(s)
        p = small_params(w, n, n, step=w // 2, rShiftPoly=[0.6 * k, 0, 0], aShiftPoly=[0, 0, 0])
        out.append((abi.ST_JOB_STRACK, p, a, b, dict(islandThresh=3) if k == 1 else None))
    s = configs._spec(480, 480, 310, dr0=1.7, drR=2e-3, da0=-2.4, daA=3e-3, gmin=0.7, gmax=0.9)
    a, b = track.synth_host(s)
    p = stw_params(64, 480, 480, 8, navg=2, step=32, rShiftPoly=[1.7, 2e-3, 0], aShiftPoly=[-2.4, 0, 3e-3])
    out.append((abi.ST_JOB_STRACKW, p, a, b, None))
    s, p = configs.cfg1(scale=8)
    a, b = track.synth_host(s)
    out.append((abi.ST_JOB_STRACK, p, a, b, None))
    return out


def _alone(dev, kind, p, a, b, cullkw):
    pid = dev.upload_pair(a, b)
    try:
        r = dev.speckle_track(pid, p) if kind == abi.ST_JOB_STRACK else dev.corr_track_fast(pid, p)
    finally:
        dev.release_pair(pid)
    extra = None
    if cullkw is not None:
        cp, extra = track.cull_for_result(**cullkw)(r)
        dev.cull("st_cull_run", cp)
    return r, extra


def _same(r1, r2):
    for f in ("offR", "offA", "corr", "type"):
        assert np.array_equal(getattr(r1, f), getattr(r2, f)), f
    assert (r1.nComplex, r1.nAmp, r1.nAmpL, r1.nFail) == (r2.nComplex, r2.nAmp, r2.nAmpL, r2.nFail)


def test_results_in_the_mix_equal_results_alone(device):
    jobs = _jobs()
    alone = [_alone(device, *j) for j in jobs]
    srv = track.Server(0, slots=3, hbm_budget=64 << 30)
    try:
        # one job on a resident pair, the others upload their own images;
        # two rounds so that later jobs arrive while earlier ones run
        rp = srv.upload_pair(jobs[2][2], jobs[2][3])
        ids = []
        for rnd in range(2):
            for k, (kind, p, a, b, cullkw) in enumerate(jobs):
                cull = track.cull_for_result(**cullkw) if cullkw is not None else None
                if k == 2:
                    ids.append((k, srv.submit(p, pair=rp, kind=kind, cull=cull)))
                else:
                    ids.append((k, srv.submit(p, images=(a, b), kind=kind, cull=cull)))
        slots = set()
        for k, jid in ids:
            r, info, extra = srv.wait(jid)
            slots.add(info.slot)
            assert info.status == abi.ST_OK and info.latencyMs >= info.runMs > 0
            assert info.patches > 0
            _same(r, alone[k][0])
            if extra is not None:
                for f in ("offRS", "offAS", "sigmaR", "sigmaA"):
                    assert np.array_equal(extra[f], alone[k][1][f]), f
        assert len(slots) > 1
        srv.release_pair(rp)
    finally:
        srv.close()


def _admission_job():
    """A strack job on a 2048 x 1024 pair with a sparse grid: its images
    (32 MB) are large next to its scratch needs."""
    s = configs._spec(2048, 1024, 321, dr0=0.4, da0=-0.3, gmin=0.6, gmax=0.9)
    a, b = track.synth_host(s)
    p = small_params(32, 2048, 1024, step=96, rShiftPoly=[0.4, 0, 0], aShiftPoly=[-0.3, 0, 0])
    return abi.ST_JOB_STRACK, p, a, b


def test_admission_by_budget(device):
    """A budget that leaves room for one job's images at a time runs the
    jobs one after another (the last one's latency covers every run); a
    job that can never fit is refused at submit."""
    kind, p, a, b = _admission_job()
    img = a.nbytes + b.nbytes
    # slot scratch is budget / (2 slots) each, so half the budget is left
    # for images: 1.5 jobs' worth
    srv = track.Server(0, slots=2, hbm_budget=3 * img)
    try:
        ids = [srv.submit(p, images=(a, b), kind=kind) for _ in range(4)]
        ref, runs, last = None, 0.0, 0.0
        for jid in ids:
            r, info, _ = srv.wait(jid)
            ref = ref or r
            _same(r, ref)
            runs += info.runMs
            last = max(last, info.latencyMs)
        assert last >= runs - 5.0
        big = np.zeros((16384, 16384), np.complex64)
        with pytest.raises(RuntimeError):
            srv.submit(p, images=(big, big), kind=kind)
    finally:
        srv.close()


def test_poll_reports_states(device):
    kind, p, a, b, _ = _jobs()[2]
    srv = track.Server(0, slots=1, hbm_budget=16 << 30)
    try:
        ids = [srv.submit(p, images=(a, b), kind=kind) for _ in range(3)]
        assert srv.poll(ids[-1]).state in (abi.ST_JOB_QUEUED, abi.ST_JOB_RUNNING)
        srv.wait(ids[-1])
        assert all(srv.poll(j).state == abi.ST_JOB_DONE for j in ids[:-1])
        assert all(srv.wait(j)[1].state == abi.ST_JOB_DONE for j in ids[:-1])
        bad = track.default_params(rStart=0, aStart=0, deltaR=8, deltaA=8, nR=4, nA=4, wR=25, wA=32, wRa=24,
                                   wAa=32)
        jid = srv.submit(bad, images=(a, b), kind=kind)
        assert srv.poll(jid).state in (abi.ST_JOB_QUEUED, abi.ST_JOB_RUNNING, abi.ST_JOB_FAILED)
        with pytest.raises(RuntimeError, match="window sizes"):
            srv.wait(jid)
        # wait() released the job: its id is no longer known
        with pytest.raises(RuntimeError, match="unknown job"):
            srv.poll(jid)
    finally:
        srv.close()


def test_cancel_queued_and_running_jobs(device):
    """On one slot: cancelling a queued job ends it at once without running
    it; cancelling the running job ends it at a batch boundary (or it
    finishes first); the jobs after them run and give the results they give
    alone."""
    kind, p, a, b, _ = _jobs()[1]
    ref, _ = _alone(device, kind, p, a, b, None)
    srv = track.Server(0, slots=1, hbm_budget=16 << 30)
    try:
        ids = [srv.submit(p, images=(a, b), kind=kind) for _ in range(5)]
        assert srv.cancel(ids[3])
        assert srv.cancel(ids[0])
        states = {}
        for j in ids:
            r, info, _ = srv.wait(j, check=False)
            states[j] = info.state
            if info.state == abi.ST_JOB_DONE:
                _same(r, ref)
            else:
                assert info.state == abi.ST_JOB_CANCELLED and info.status == abi.ST_ERR_CANCELLED
        assert states[ids[3]] == abi.ST_JOB_CANCELLED
        assert states[ids[0]] in (abi.ST_JOB_CANCELLED, abi.ST_JOB_DONE)
        assert all(states[j] == abi.ST_JOB_DONE for j in (ids[1], ids[2], ids[4]))
        # an ended job cannot be cancelled; the server still runs new work
        jid = srv.submit(p, images=(a, b), kind=kind)
        r, info, _ = srv.wait(jid)
        _same(r, ref)
        assert not srv.cancel(jid)
    finally:
        srv.close()


def test_pair_release_waits_for_its_jobs_and_refuses_new_ones(device):
    kind, p, a, b, _ = _jobs()[0]
    ref, _ = _alone(device, kind, p, a, b, None)
    srv = track.Server(0, slots=2, hbm_budget=16 << 30)
    try:
        rp = srv.upload_pair(a, b)
        ids = [srv.submit(p, pair=rp, kind=kind) for _ in range(4)]
        srv.release_pair(rp)  # returns once the four jobs have ended
        assert all(srv.poll(j).state == abi.ST_JOB_DONE for j in ids)
        for j in ids:
            _same(srv.wait(j)[0], ref)
        with pytest.raises(RuntimeError):
            srv.submit(p, pair=rp, kind=kind)
        with pytest.raises(RuntimeError):
            srv.release_pair(rp)
    finally:
        srv.close()


def test_batched_jobs_equal_jobs_alone(device):
    """Many small fused-path jobs submitted at once are taken in batches by
    the slots (one fused launch per window class per batch); with batching
    off the same jobs run one by one.  Every job gives the same grids both
    ways and alone."""
    jobs = []
    for k in range(12):
        w = (16, 32, 64)[k % 3]
        n = 256 + 32 * (k % 4)
        s = configs._spec(n, n, 700 + k, dr0=0.3 * (k % 5) - 0.6, da0=0.2, gmin=0.5, gmax=0.9)
        jobs.append((s, small_params(w, n, n, step=w // 2, rShiftPoly=[0.3 * (k % 5) - 0.6, 0, 0],
                                     aShiftPoly=[0.2, 0, 0])))
    imgs = [track.synth_host(s) for s, _ in jobs]
    alone = [_alone(device, abi.ST_JOB_STRACK, p, a, b, None)[0] for (_, p), (a, b) in zip(jobs, imgs)]
    for batch in (8, 1):
        srv = track.Server(0, slots=2, hbm_budget=32 << 30)
        try:
            srv.set_batch(batch)
            pids = [srv.upload_pair(a, b) for a, b in imgs]
            ids = [srv.submit(p, pair=pid) for (_, p), pid in zip(jobs, pids)]
            for k, jid in enumerate(ids):
                r, info, _ = srv.wait(jid)
                assert info.patches > 0
                _same(r, alone[k])
        finally:
            srv.close()
