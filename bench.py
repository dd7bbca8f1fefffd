"""Benchmark: patch correlations/s and job latency of the strack hot path
on one MI355X.

Workload (default): the full-scene single pair -- 20000 range x 14000
azimuth complex64 per image, synthetic speckle with log-normal
backscatter, glacier-ridge offset field (12 px azimuth / 4 px range peak),
coherence 0.3-0.9, a zero-filled block of 10% of the scene; strack with
complex matching plus amplitude retries, 64 x 64 windows on a 24 px grid.
The pair is generated directly in HBM and stays resident; one step is one
full tracking job over the whole offset grid.

Prints one JSON line (see the repository contract): value = patch
correlations per second, with roofline of the dominant kernel (live HIP
event timings), p50/p99 job latency and the CPU restatement timed on a
bounded strip of the same scene.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def flops_complex_patch(w):
    """Algorithmic FP32 flops of one complex patch correlation: 5 N log2 N
    per complex FFT (two w^2 forward, one (2w)^2 inverse), 6 per complex
    multiply of the cross spectrum, 4 per |z| of the surface, plus the
    8x8 -> 48x48 pruned oversampling."""
    import math
    f_os = 5 * 64 * 6 + 8 * 8 * 48 * 8 + 8 * 48 * 48 * 8 + 4 * 48 * 48
    return (10 * w * w * math.log2(w * w) + 6 * w * w + 20 * w * w * math.log2(4 * w * w)
            + 16 * w * w + f_os)


def flops_amp_patch(wa):
    """Amplitude try: forward wa^2 x2, inverse (2wa)^2 x2 (upsampling),
    forward (2wa)^2 x2, inverse (2wa)^2, detection and cross products,
    plus the oversampling."""
    import math
    n1, n2 = wa * wa, 4 * wa * wa
    f_os = 5 * 64 * 6 + 8 * 8 * 48 * 8 + 8 * 48 * 48 * 8 + 4 * 48 * 48
    return (2 * 5 * n1 * math.log2(n1) + 5 * 5 * n2 * math.log2(n2) + 6 * n2 + 8 * n2 + f_os)


TRAFFIC_PROFILE = os.path.join(HERE, "profiles", "k_cpx_fused_fullscene_hbm_traffic.md")


def measured_traffic():
    """HBM bytes per launch of the fused complex-match kernel, from the
    committed FETCH_SIZE / WRITE_SIZE profile of this workload (None when
    absent)."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for ln in f:
                if ln.startswith("traffic_bytes_per_launch:"):
                    return int(float(ln.split(":")[1]))
    except OSError:
        pass
    return None


def calibrate_copy(nbytes=1 << 30, reps=5):
    """A plain device-to-device copy of known size, for scaling the HBM
    byte counters (FETCH_SIZE / WRITE_SIZE) of a profiled run."""
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda:0").fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / reps
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "GB_per_s": 2 * nbytes / dt / 1e9}))


def cpu_baseline(spec, par, dev, pid, budget_s=15.0):
    """Time the FP32 CPU restatement (single thread) on a strip of grid
    rows of the same resident scene."""
    from oracle.pyoracle import Oracle
    from specklesource_amd import track
    o = Oracle("f32")
    rows = 1
    total_pts = 0
    total_t = 0.0
    total_corr = 0
    i0 = par.nA // 3
    while total_t < budget_s and i0 + rows <= par.nA:
        sub = track.default_params()
        for f, _ in par._fields_:
            setattr(sub, f, getattr(par, f))
        sub.nA = rows
        sub.aStart = par.aStart + i0 * par.deltaA
        a0 = max(0, sub.aStart - 2 * par.wA - 64)
        a1 = min(spec.na, sub.aStart + (rows - 1) * par.deltaA + 2 * par.wA + 64)
        img1, img2 = dev.download_pair(pid, a0, a1 - a0, spec.nr)
        sub.aStart -= a0
        sub.aShiftPoly[0] = par.aShiftPoly[0] + par.aShiftPoly[2] * a0
        sub.rShiftPoly[0] = par.rShiftPoly[0] + par.rShiftPoly[2] * a0
        t = time.perf_counter()
        r = o.speckle_track(sub, img1, img2, diag=True)
        dt = time.perf_counter() - t
        total_t += dt
        total_pts += sub.nR * sub.nA
        total_corr += int(sum((r.diag["peakI"][k] > -2000).sum() for k in range(3)))
        i0 += rows
        rows = min(4, rows * 2)
    return {"value": total_corr / total_t, "unit": "patch correlations/s", "cores": 1, "kind": "port",
            "sample": "%d grid points (%d consecutive grid rows of the same scene), %.1f s single-thread"
                      % (total_pts, total_pts // par.nR, total_t)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=1, help="divide the scene sides (testing only)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a 1 GiB device copy (for scaling HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(sharding design in DESIGN.md); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return

    from specklesource_amd import configs, track
    spec, par = configs.cfg2(scale=args.scale)
    dev = track.Device(0)
    t0 = time.perf_counter()
    pid = dev.synth_pair(spec)
    gen_s = time.perf_counter() - t0
    print("generated %d x %d pair in HBM in %.1f s" % (spec.nr, spec.na, gen_s), file=sys.stderr, flush=True)

    for k in range(args.warmup):
        dev.speckle_track(pid, par)
        print("warmup %d done" % k, file=sys.stderr, flush=True)
    dev.set_profiling(True)
    lat = []
    ncorr = 0
    kern = {}
    stages = np.zeros(4)
    res = None
    t_start = time.perf_counter()
    for k in range(args.steps):
        t = time.perf_counter()
        res = dev.speckle_track(pid, par)
        lat.append(time.perf_counter() - t)
        st = dev.last_stats()
        ncorr += sum(st.nCorr[i] for i in range(3))
        stages += np.array([st.stageMs[i] for i in range(4)])
        for i in range(8):
            nm = st.kernelNames[i].value.decode()
            if nm:
                ms, n = kern.get(nm, (0.0, 0))
                kern[nm] = (ms + st.kernelMs[i], n + st.kernelLaunches[i])
        print("step %d: %.3f s" % (k, lat[-1]), file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t_start
    dev.set_profiling(False)
    value = ncorr / elapsed
    st = dev.last_stats()
    ncpx = st.nCorr[0]
    # roofline of the dominant kernel: the fused complex-match kernel does
    # every complex patch correlation of a job in one launch, so its
    # algorithmic flops per launch are F_c(w) x complex correlations per job,
    # over its average launch duration (HIP events on the library's stream)
    dom = max(kern.items(), key=lambda kv: kv[1][0]) if kern else ("none", (0, 0))
    if "k_cpx_fused" in kern and kern["k_cpx_fused"][1] > 0:
        kname = "k_cpx_fused"
        launch_ms = kern[kname][0] / kern[kname][1]
        per_launch = flops_complex_patch(par.wR) * ncpx * args.steps / kern[kname][1]
    else:
        kname = "complex-match stage"
        launch_ms = stages[1] / args.steps
        per_launch = flops_complex_patch(par.wR) * ncpx
    achieved = per_launch / (launch_ms * 1e-3) / 1e12 if launch_ms > 0 else 0.0
    line = {
        "metric": "patch correlations/sec + p99 job latency, 1 MI355X, at <=0.02 px offset error",
        "value": value,
        "unit": "patch correlations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (fp64 sub-pixel refinement)",
        "data": "synthetic (generated in HBM)",
        "config": {"workload": "full-scene single pair, strack complex + amplitude, 64x64 windows, 24 px grid",
                   "scene": [spec.nr, spec.na], "grid": [par.nR, par.nA], "window": par.wR},
        "latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)), "p99": 1e3 * float(np.percentile(lat, 99))},
        "work": {"patch_correlations_per_job": ncorr // args.steps, "complex": int(ncpx),
                 "amp_small": int(st.nCorr[1]), "amp_large": int(st.nCorr[2]),
                 "matches": {"complex": res.nComplex, "amp": res.nAmp, "ampL": res.nAmpL, "fail": res.nFail}},
        "stage_ms": {"geometry": stages[0] / args.steps, "complex": stages[1] / args.steps,
                     "amp_small": stages[2] / args.steps, "amp_large": stages[3] / args.steps},
        "kernels_ms_per_step": {k: v[0] / args.steps for k, v in kern.items()},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": measured_traffic(),
                     "kernel": kname, "launch_ms": launch_ms, "flops_per_launch": per_launch,
                     "algorithmic_bytes_per_launch": 2 * 8 * par.wR * par.wR * ncpx,
                     "dominant_kernel": dom[0]},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(spec, par, dev, pid)
    dev.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
