"""A served stream shaped like the serving-plus-culling workload, scaled
down: seeded Poisson arrivals of tracking jobs drawn from the other four
workloads (amplitude-only scenes, complex scenes with 64 px windows,
uneven pairs with 16-128 px windows, strackw with 384 / 512 px windows),
each culled by cullst (box 5 or 9, islands below 20) -- strack jobs on the
device, straight from tracking -- and cull-only jobs on the tracked grids
with 5 % of the points moved by uniform +/-30 px.

Every job is checked against the oracle, not only against itself alone:
* the tracking (run alone, before the stream) agrees with the FP32 and
  FP64 oracles to the strack / strackw parity bars;
* the served job's culled grids equal the oracle's cullst chain run on
  those tracked grids (culled offsets and types bit for bit, sigma and
  smoothed offsets to 2e-6 relative);
* each cull-only job equals the oracle's chain on the same input.
Job latencies (p50 / p99) are printed."""
import time

import numpy as np
import pytest

from specklesource_amd import abi, configs, track
from tests.helpers import compare_strack, cull_params
from tests.test_gpu_strackw import compare as compare_w

pytestmark = pytest.mark.gpu

RTOL = 2e-6  # sigma / smoothed offsets: same FP64 sums, libm-level rounding
CULL_FIELDS = ("offR", "offA", "type")
CULL_CLOSE = ("offRS", "offAS", "sigmaR", "sigmaA")


def _alone(device, job, a, b):
    pid = device.upload_pair(a, b)
    try:
        if job["kind"] == "strack":
            return device.speckle_track(pid, job["par"], diag=True)
        return device.corr_track_fast(pid, job["par"], diag=True)
    finally:
        device.release_pair(pid)


def _oracle_cull(oracle32, offR, offA, corr, typ, kw):
    cp, arrs = cull_params(offR, offA, corr, typ, **kw)
    oracle32.cull("st_cull_run", cp)
    return arrs


def _same_cull(got, want):
    for f in CULL_FIELDS:
        assert np.array_equal(got[f], want[f]), f
    for f in CULL_CLOSE:
        a, b = got[f], want[f]
        ok = np.isclose(a, b, rtol=RTOL, atol=1e-6) | ((a < -1e9) & (b < -1e9))
        assert ok.all(), f


def test_poisson_stream_with_cull_matches_the_oracle(device, oracle32, oracle64):
    # seed 24 draws every workload and every window size (16-512)
    jobs = configs.cfg5_stream(n_jobs=10, scale=8, seed=24, rate=20.0)
    data = []
    for k, j in enumerate(jobs):
        a, b = track.synth_host(j["spec"])
        g = _alone(device, j, a, b)
        if j["kind"] == "strack":
            o32 = oracle32.speckle_track(j["par"], a, b, diag=True)
            o64 = oracle64.speckle_track(j["par"], a, b, diag=True)
            rep, probs = compare_strack(j["par"], g, o32, o64)
        else:
            o32 = oracle32.corr_track_fast(j["par"], a, b, diag=True)
            o64 = oracle64.corr_track_fast(j["par"], a, b, diag=True)
            rep, probs = compare_w(g, o32, o64)
            np.testing.assert_array_equal(g.type, o32.type)
        print("job %d (from workload %d, %s, w=%d): %s" % (k, j["src"], j["kind"], j["par"].wR, rep), flush=True)
        assert not probs, probs[:10]
        if "rms_r_vs_f64" in rep:
            assert rep["rms_r_vs_f64"] <= 0.02 and rep["rms_a_vs_f64"] <= 0.02
        want = _oracle_cull(oracle32, g.offR, g.offA, g.corr, g.type, j["cull"])
        r_out, a_out = configs.inject_outliers(g.offR, g.offA, 0.05, seed=100 + k)
        want_out = _oracle_cull(oracle32, r_out, a_out, g.corr, g.type, j["cull"])
        data.append((a, b, want, (r_out, a_out, g.corr, g.type), want_out))

    srv = track.Server(0, slots=4, hbm_budget=200 << 30)
    try:
        t0 = time.perf_counter()
        ids = []
        for k, j in enumerate(jobs):
            delay = t0 + j["arrival"] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            a, b, _, (ro, ao, c, ty), _ = data[k]
            kind = abi.ST_JOB_STRACK if j["kind"] == "strack" else abi.ST_JOB_STRACKW
            tid = srv.submit(j["par"], images=(a, b), kind=kind, cull=track.cull_for_result(**j["cull"]))
            cp, arrs = cull_params(ro, ao, c, ty, **j["cull"])
            cid = srv.submit_cull(cp, arrs)
            ids.append((k, tid, cid))
        lat = []
        for k, tid, cid in ids:
            r, info, extra = srv.wait(tid)
            lat.append(info.latencyMs)
            got = dict(offR=r.offR, offA=r.offA, type=r.type, **extra)
            _same_cull(got, data[k][2])
            _, cinfo, carrs = srv.wait(cid)
            lat.append(cinfo.latencyMs)
            _same_cull(carrs, data[k][4])
        print("stream of %d jobs: latency p50 %.1f ms, p99 %.1f ms"
              % (len(lat), np.percentile(lat, 50), np.percentile(lat, 99)))
    finally:
        srv.close()
