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


def flops_window_sums(w):
    """Flops of the window sums the fused kernel forms for w <= 64
    (getPatches' statistics and the Doppler lag sum): per sample of each of
    the two windows, |x|^2 (3), the running sum (2) and the lag product
    accumulated (8).  At w = 128 a separate kernel forms them."""
    return 26 * w * w if w <= 64 else 0


def flops_amp_patch(wa):
    """Amplitude try: forward wa^2 x2, inverse (2wa)^2 x2 (upsampling),
    forward (2wa)^2 x2, inverse (2wa)^2, detection and cross products,
    plus the oversampling."""
    import math
    n1, n2 = wa * wa, 4 * wa * wa
    f_os = 5 * 64 * 6 + 8 * 8 * 48 * 8 + 8 * 48 * 48 * 8 + 4 * 48 * 48
    return (2 * 5 * n1 * math.log2(n1) + 5 * 5 * n2 * math.log2(n2) + 6 * n2 + 8 * n2 + f_os)


TRAFFIC_PROFILE = os.path.join(HERE, "profiles", "k_cpx_fused_fullscene_hbm_traffic.md")
SERVED_TRAFFIC_PROFILE = os.path.join(HERE, "profiles", "k_cpx_fused_served_hbm_traffic.md")


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


def served_traffic():
    """HBM bytes per fused launch (k_cpx_fused plus its k_cpx_refine) of the
    served workload's one-context pass, from the committed FETCH_SIZE /
    WRITE_SIZE profile (None when absent)."""
    try:
        with open(SERVED_TRAFFIC_PROFILE) as f:
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


def flops_strackw_patch(wa):
    """Algorithmic FP32 flops of one strackw patch, counted from the
    reference's transform list: forward wa^2 x2, inverse (2wa)^2 x2,
    forward (2wa)^2 x2, inverse (2wa)^2, the 336^2 inverse of the
    oversampling, the cross multiply and the |z| / detection passes."""
    import math
    n1, n2, n3 = wa * wa, 4 * wa * wa, 336 * 336
    return (2 * 5 * n1 * math.log2(n1) + 5 * 5 * n2 * math.log2(n2) + 5 * n3 * math.log2(n3)
            + 6 * n2 + 3 * 4 * n2)


def bench_strackw(args):
    """Sheared-flow strackw workload: 8192 x 8192 pair, coherence 0.15-0.4,
    shear to +/-120 px, 5x5 boxcar, chip 128 (wRa 384) or 256 (wRa 512),
    edge pad 123, initial-offset grid at 1/8 resolution with clipped
    2.5 px Gaussian error."""
    import ctypes as C
    from specklesource_amd import abi, configs, track
    spec, par = configs.cfg4(chip=args.chip, scale=args.scale)
    gr, ga, (r0, a0, dR, dA) = configs.initial_grid(spec, par)
    par.polyShift = 0
    par.initDr = abi.ptr(gr, C.c_float)
    par.initDa = abi.ptr(ga, C.c_float)
    par.initNa, par.initNr = gr.shape
    par.initR0, par.initA0, par.initDeltaR, par.initDeltaA = r0, a0, dR, dA
    dev = track.Device(0)
    pid = dev.synth_pair(spec)
    for k in range(args.warmup):
        dev.corr_track_fast(pid, par)
        print("warmup %d done" % k, file=sys.stderr, flush=True)
    dev.set_profiling(True)
    lat, kern, npatch, res = [], {}, 0, None
    t0 = time.perf_counter()
    for k in range(args.steps):
        t = time.perf_counter()
        res = dev.corr_track_fast(pid, par)
        lat.append(time.perf_counter() - t)
        st = dev.last_stats()
        npatch += st.nCorr[1]
        for i in range(len(st.kernelNames)):
            nm = st.kernelNames[i].value.decode()
            if nm:
                ms, n = kern.get(nm, (0.0, 0))
                kern[nm] = (ms + st.kernelMs[i], n + st.kernelLaunches[i])
        print("step %d: %.3f s" % (k, lat[-1]), file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0
    dev.set_profiling(False)
    per = npatch // args.steps
    fl = flops_strackw_patch(par.wRa)
    busy_ms = sum(v[0] for v in kern.values()) / args.steps
    dom = max(kern.items(), key=lambda kv: kv[1][0])
    line = {
        "metric": "strackw patch correlations/sec, 1 MI355X",
        "value": npatch / elapsed, "unit": "patch correlations/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32 (fp64 window sums, NCC and refinement)",
        "data": "synthetic (generated in HBM)",
        "config": {"workload": "strackw sheared flow, chip %d (wRa %d), edge pad 123, 5x5 boxcar"
                               % (args.chip, par.wRa), "scene": [spec.nr, spec.na], "grid": [par.nR, par.nA]},
        "latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)), "p99": 1e3 * float(np.percentile(lat, 99))},
        "work": {"patches_per_job": per, "matches": res.nAmp, "fail": res.nFail},
        "kernels_ms_per_step": {k: v[0] / args.steps for k, v in kern.items()},
        "roofline": {"bound": "valu", "achieved": fl * per / (busy_ms * 1e-3) / 1e12 if busy_ms > 0 else 0.0,
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "traffic": None,
                     "kernel": "all strackw kernels (algorithmic flops of the reference transform list over "
                               "the summed kernel time)", "dominant_kernel": dom[0]},
    }
    line["roofline"]["frac"] = line["roofline"]["achieved"] / PEAK_FP32_TFLOPS
    dev.close()
    print(json.dumps(line))


def bench_batched(args):
    """Batched uneven workload: 64 pairs with sides drawn from 1024-8192 px,
    windows from {16, 32, 64, 128} on a w/2 grid, affine offsets within
    +/-20 px, all resident in HBM and submitted at once to the serving
    layer (4 slots); one step is the whole set of 64 jobs."""
    from specklesource_amd import abi, configs, track
    prec = abi.ST_PREC_MFMA if args.precision == "mfma" else abi.ST_PREC_FP32
    pairs = configs.cfg3_pairs(n_pairs=args.pairs, scale=args.scale)
    if args.windows:
        keep = {int(x) for x in args.windows.split(",")}
        pairs = [(s, p) for s, p in pairs if p.wR in keep]
    srv = track.Server(0, slots=args.slots, hbm_budget=200 << 30)
    t0 = time.perf_counter()
    pids = [srv.synth_pair(s) for s, _ in pairs]
    print("generated %d pairs in HBM in %.1f s" % (len(pids), time.perf_counter() - t0), file=sys.stderr, flush=True)
    for _, p in pairs:
        p.precision = prec

    def run_all():
        t = time.perf_counter()
        ids = [srv.submit(p, pair=pid) for (_, p), pid in zip(pairs, pids)]
        infos = [srv.wait(j)[1] for j in ids]
        return time.perf_counter() - t, infos

    for k in range(args.warmup):
        run_all()
        print("warmup %d done" % k, file=sys.stderr, flush=True)
    walls, lats, patches = [], [], 0
    for k in range(args.steps):
        wall, infos = run_all()
        walls.append(wall)
        lats += [i.latencyMs for i in infos]
        patches += sum(i.patches for i in infos)
        print("step %d: %.3f s" % (k, wall), file=sys.stderr, flush=True)
    srv.close()
    ws = {}
    for _, p in pairs:
        ws[p.wR] = ws.get(p.wR, 0) + 1
    line = {
        "metric": "patch correlations/sec + p99 job latency, 1 MI355X (batched uneven pairs)",
        "value": patches / sum(walls), "unit": "patch correlations/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * sum(walls) / len(walls), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32" if prec == abi.ST_PREC_FP32 else "bf16 split hi+lo, fp32 accumulate",
        "data": "synthetic (generated in HBM)",
        "config": {"workload": "%d pairs 1024-8192 px, w in {16,32,64,128}, grid w/2, served on %d slots"
                               % (len(pairs), args.slots), "windows": ws, "precision": args.precision},
        "latency_ms": {"p50": float(np.percentile(lats, 50)), "p99": float(np.percentile(lats, 99))},
        "work": {"patch_correlations_per_step": patches // args.steps},
    }
    print(json.dumps(line))


def flops_mfma_patch(w):
    """Matrix-core work of one complex patch on the split-bf16 DFT path: six
    w x w complex 2-D DFT contractions (two forward, four polyphase
    inverses), 16 w^3 flops each as two complex w x w x w products, three
    products for the hi/lo split."""
    return 6 * 16 * w ** 3 * 3


PEAK_BF16_TFLOPS = 2500.0


def _kernel_table(dev, table):
    st = dev.last_stats()
    for i in range(len(st.kernelNames)):
        nm = st.kernelNames[i].value.decode()
        if nm:
            ms, n = table.get(nm, (0.0, 0))
            table[nm] = (ms + st.kernelMs[i], n + st.kernelLaunches[i])
    return st


def served_roofline(pairs, precision):
    """Dominant-kernel roofline of the served workload, measured on one
    context (no other slot running, so each launch's HIP-event duration is
    its own): the same jobs run one after another with kernel timing on,
    and the complex-match kernel's algorithmic work over all of them is
    divided by its summed launch time."""
    from specklesource_amd import abi, track
    dev = track.Device(0)
    pids = [dev.synth_pair(s) for s, _ in pairs]
    dev.set_profiling(True)
    table, work, refine = {}, {}, {}
    for k, ((s, p), pid) in enumerate(zip(pairs, pids)):
        q = abi.TrackParams.from_buffer_copy(p)
        q.precision = precision
        dev.speckle_track(pid, q)
        print("roofline pass: job %d of %d done" % (k + 1, len(pairs)), file=sys.stderr, flush=True)
        st = _kernel_table(dev, table)
        names = [st.kernelNames[i].value.decode() for i in range(len(st.kernelNames))]
        kern = "k_cpx_mfma" if "k_cpx_mfma" in names else ("k_cpx_fused" if "k_cpx_fused" in names else "staged")
        fl = flops_mfma_patch(p.wR) if kern == "k_cpx_mfma" else flops_complex_patch(p.wR) + flops_window_sums(p.wR)
        f, n, b = work.get(kern, (0.0, 0, 0))
        work[kern] = (f + fl * st.nCorr[0], n + st.nCorr[0], b + 16 * p.wR * p.wR * st.nCorr[0])
        # this call's refinement-kernel time, charged to its own match kernel
        if "k_cpx_refine" in names:
            k = names.index("k_cpx_refine")
            rms, rn = refine.get(kern, (0.0, 0))
            refine[kern] = (rms + st.kernelMs[k], rn + st.kernelLaunches[k])
    dev.set_profiling(False)
    for pid in pids:
        dev.release_pair(pid)
    dev.close()
    dom = max((k for k in table if k in ("k_cpx_fused", "k_cpx_mfma")), key=lambda k: table[k][0])
    ms, launches = table[dom]
    flops, patches, byts = work[dom]
    mfma = dom == "k_cpx_mfma"
    label = dom
    by_kernel = {dom: table[dom][0] / table[dom][1]}
    traffic = None
    if dom == "k_cpx_fused" and dom in refine:
        # the fused path's sub-pixel refinement runs in its own kernel right
        # after it: the complex match's flops are divided by both
        ms += refine[dom][0]
        by_kernel["k_cpx_refine"] = refine[dom][0] / refine[dom][1]
        label = "k_cpx_fused + k_cpx_refine"
        traffic = served_traffic()
    achieved = flops / (ms * 1e-3) / 1e12
    peak = PEAK_BF16_TFLOPS if mfma else PEAK_FP32_TFLOPS
    return {"bound": "mfma" if mfma else "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
            "frac": achieved / peak, "traffic": traffic, "kernel": label, "launches": launches,
            "launch_ms_avg": ms / launches, "launch_ms_avg_by_kernel": by_kernel, "patches": patches,
            "algorithmic_bytes_per_launch": byts / launches,
            "flops_total": flops, "algorithmic_bytes_total": byts,
            "kernel_ms_per_pass": {k: v[0] for k, v in sorted(table.items(), key=lambda kv: -kv[1][0])},
            "measured": "one context, the same %d jobs in sequence, HIP events on the library stream" % len(pairs)}


def _oracle_strip(job):
    """CPU baseline worker: the FP32 oracle on a strip of grid rows (the
    parameter block comes as bytes: it holds no pointers for this
    workload)."""
    from oracle.pyoracle import Oracle
    from specklesource_amd import abi
    par, img1, img2 = job
    if isinstance(par, (bytes, bytearray)):
        par = abi.TrackParams.from_buffer_copy(par)
    t = time.perf_counter()
    r = Oracle("f32").speckle_track(par, img1, img2, diag=True)
    dt = time.perf_counter() - t
    return dt, int(sum((r.diag["peakI"][k] > -2000).sum() for k in range(3)))


def _strips(pairs, dev, pids, rows_per_strip=2, per_pair=2):
    """Bounded sample of the served workload: `per_pair` strips of grid rows
    from every pair, with their images read back from HBM."""
    from specklesource_amd import track
    out = []
    for (spec, par), pid in zip(pairs, pids):
        for k in range(per_pair):
            i0 = (par.nA * (k + 1)) // (per_pair + 1)
            rows = min(rows_per_strip, par.nA - i0)
            if rows <= 0:
                continue
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
            out.append((sub, img1, img2))
    return out


def cpu_cores():
    """Host cores this process may use, at most 16 (one GPU's share of the
    box)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def served_cpu_and_accuracy(pairs, budget_s=12.0):
    """CPU baseline (the FP32 oracle restatement, one core and all cores)
    and accuracy of both GPU precisions on a bounded sample of the served
    workload: two strips of two grid rows from every pair."""
    import concurrent.futures as cf
    from oracle.pyoracle import Oracle
    from specklesource_amd import abi, track
    from tests.helpers import compare_strack
    dev = track.Device(0)
    pids = [dev.synth_pair(s) for s, _ in pairs]
    jobs = _strips(pairs, dev, pids)
    for pid in pids:
        dev.release_pair(pid)
    # one core, in order, until the budget is spent
    t1 = n1 = 0
    for j in jobs:
        dt, n = _oracle_strip(j)
        t1 += dt
        n1 += n
        if t1 > budget_s:
            break
    cores = cpu_cores()
    t = time.perf_counter()
    nall = 0
    with cf.ProcessPoolExecutor(max_workers=cores) as ex:
        for dt, n in ex.map(_oracle_strip, [(bytes(p), a, b) for p, a, b in jobs]):
            nall += n
    tall = time.perf_counter() - t
    base1 = {"value": n1 / t1, "unit": "patch correlations/s", "cores": 1, "kind": "port",
             "sample": "%d patch correlations from strips of 2 grid rows of the 64 pairs, %.1f s single-thread"
                       % (n1, t1)}
    baseN = {"value": nall / tall, "unit": "patch correlations/s", "cores": cores, "kind": "port",
             "sample": "%d patch correlations (2 strips of 2 grid rows from each of the 64 pairs) in %d "
                       "processes, %.1f s wall" % (nall, cores, tall)}
    # accuracy of both precisions on the same strips
    o32, o64 = Oracle("f32"), Oracle("f64")
    acc = {}
    for prec_name, prec in (("fp32", abi.ST_PREC_FP32), ("mfma", abi.ST_PREC_MFMA)):
        tot = agree = ties = unexpl = 0
        kinds = {}
        er, ea, cr, sr = [], [], [], []
        for sub, img1, img2 in jobs:
            q = abi.TrackParams.from_buffer_copy(sub)
            q.precision = prec
            pid = dev.upload_pair(img1, img2)
            g = dev.speckle_track(pid, q, diag=True)
            dev.release_pair(pid)
            r32 = o32.speckle_track(sub, img1, img2, diag=True)
            r64 = o64.speckle_track(sub, img1, img2, diag=True)
            rep, probs = compare_strack(q, g, r32, r64, lm_rtol=2e-2 if prec else 2e-3,
                                        tie=5e-4 if prec else 2e-5, thresh_tol=1e-3 if prec else 1e-4)
            tot += rep["attempts"]
            agree += round(rep["peak_agree"] * rep["attempts"])
            ties += rep["near_ties"]
            unexpl += len(probs)
            for pr in probs:
                kinds[str(pr[0])] = kinds.get(str(pr[0]), 0) + 1
            both = (g.type > 0) & (r64.type == g.type)
            er += list((g.offR - r64.offR)[both].astype(np.float64))
            ea += list((g.offA - r64.offA)[both].astype(np.float64))
            cr += list((np.abs(g.corr.astype(np.float64) - r64.corr) / np.maximum(np.abs(r64.corr), 1e-12))[both])
            for t in range(3):
                ran = (g.diag["peakI"][t] > -1000) & (r64.diag["peakI"][t] == g.diag["peakI"][t])
                if ran.any():
                    a, b = g.diag["snr"][t][ran].astype(np.float64), r64.diag["snr"][t][ran].astype(np.float64)
                    sr += list(np.abs(a - b) / np.maximum(np.abs(b), 1e-12))
        acc[prec_name] = {
            "sample_points": int(sum(j[0].nR * j[0].nA for j in jobs)), "attempts": tot,
            "integer_peak_agreement_vs_fp32_oracle": agree / max(tot, 1), "near_ties": ties,
            "unexplained_mismatches": unexpl, "mismatch_kinds": kinds,
            "rms_offset_px_vs_fp64": [float(np.sqrt(np.mean(np.square(er)))) if er else None,
                                      float(np.sqrt(np.mean(np.square(ea)))) if ea else None],
            "corr_relerr_median_vs_fp64": float(np.median(cr)) if cr else None,
            "corr_relerr_max_vs_fp64": float(np.max(cr)) if cr else None,
            "snr_relerr_median_vs_fp64": float(np.median(sr)) if sr else None,
            "snr_relerr_max_vs_fp64": float(np.max(sr)) if sr else None}
    dev.close()
    return base1, baseN, acc


def served_passes(srv, pairs, pids, precision, warmup, steps, label):
    from specklesource_amd import abi

    def run_all():
        t = time.perf_counter()
        ids = []
        for (_, p), pid in zip(pairs, pids):
            q = abi.TrackParams.from_buffer_copy(p)
            q.precision = precision
            ids.append(srv.submit(q, pair=pid))
        infos = [srv.wait(j)[1] for j in ids]
        return time.perf_counter() - t, infos

    for k in range(warmup):
        run_all()
        print("%s warmup %d done" % (label, k), file=sys.stderr, flush=True)
    walls, lats, patches = [], [], 0
    for k in range(steps):
        wall, infos = run_all()
        walls.append(wall)
        lats += [i.latencyMs for i in infos]
        patches += sum(i.patches for i in infos)
        print("%s step %d: %.3f s" % (label, k, wall), file=sys.stderr, flush=True)
    return patches / sum(walls), walls, lats


def bench_served(args):
    """The batched uneven workload through the serving layer: 64 pairs with
    sides drawn from 1024-8192 px, windows from {16, 32, 64, 128} on a w/2
    grid, affine offsets within +/-20 px, resident in HBM; one step submits
    the 64 jobs to a 4-slot server and waits for all of them.  Both the FP32
    path and the split-bf16 matrix-core path are timed; `value` is the run
    with each window's tuned precision (ST_PREC_AUTO)."""
    from specklesource_amd import abi, configs, track
    pairs = configs.cfg3_pairs(n_pairs=args.pairs, scale=args.scale)
    srv = track.Server(0, slots=args.slots, hbm_budget=200 << 30)
    srv.set_batch(args.batch)
    t0 = time.perf_counter()
    pids = [srv.synth_pair(s) for s, _ in pairs]
    print("generated %d pairs in HBM in %.1f s" % (len(pids), time.perf_counter() - t0), file=sys.stderr, flush=True)
    v32, w32, l32 = served_passes(srv, pairs, pids, abi.ST_PREC_FP32, args.warmup, args.steps, "fp32")
    vmf, wmf, lmf = served_passes(srv, pairs, pids, abi.ST_PREC_MFMA, args.warmup, args.steps, "mfma")
    vau, wau, lau = served_passes(srv, pairs, pids, abi.ST_PREC_AUTO, args.warmup, args.steps, "auto")
    srv.close()
    dev = track.Device(0)
    tune = {w: dev.tune(w) for w in (16, 32, 64, 128)}
    dev.close()
    ws = {}
    for _, p in pairs:
        ws[p.wR] = ws.get(p.wR, 0) + 1
    line = {
        "metric": "patch correlations/sec + p99 job latency, 1 MI355X, at <=0.02 px offset error",
        "value": vau, "unit": "patch correlations/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * sum(wau) / len(wau), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": "per window as tuned (%s; fp64 sub-pixel refinement and normalisation sums)"
                 % ", ".join("%d: %s" % (w, "fp32" if not t or t[0] == 0 else "bf16 split")
                             for w, t in tune.items()),
        "data": "synthetic (generated in HBM)",
        "config": {"workload": "%d uneven pairs 1024-8192 px, w in {16,32,64,128}, grid w/2, search w/4, "
                               "served on %d slots" % (len(pairs), args.slots), "windows": ws,
                   "batch": args.batch},
        "latency_ms": {"p50": float(np.percentile(lau, 50)), "p99": float(np.percentile(lau, 99))},
        "tune": {str(w): (None if t is None else {"precision": ["fp32", "mfma"][t[0]], "wg_per_cu": t[1],
                                                  "staged_batch": t[2], "us_per_patch": t[3]})
                 for w, t in tune.items()},
        "precisions": {
            "auto": {"value": vau, "latency_ms": {"p50": float(np.percentile(lau, 50)),
                                                  "p99": float(np.percentile(lau, 99))}},
            "fp32": {"value": v32, "latency_ms": {"p50": float(np.percentile(l32, 50)),
                                                  "p99": float(np.percentile(l32, 99))}},
            "mfma": {"value": vmf, "latency_ms": {"p50": float(np.percentile(lmf, 50)),
                                                  "p99": float(np.percentile(lmf, 99))},
                     "dtype": "bf16 hi+lo split products, fp32 accumulate (w = 16, 32; larger windows fp32)"},
            "mfma_over_fp32": vmf / v32, "target_ratio": 3.0},
    }
    if not args.no_roofline:
        line["roofline"] = served_roofline(pairs, abi.ST_PREC_FP32)
        line["roofline_mfma"] = served_roofline(pairs, abi.ST_PREC_MFMA)
    if not args.no_cpu:
        line["cpu_baseline"], line["cpu_baseline_all_cores"], line["accuracy"] = served_cpu_and_accuracy(pairs)
    print(json.dumps(line))


def bench_cull(args):
    """cullst on the full-scene offset grid (821 x 571 points) with 5%
    uniform +/-30 px outliers: three median passes, plane-fit sigma,
    smoothing, island removal (threshold 20)."""
    from specklesource_amd import configs, track
    offR, offA, corr, typ = configs.make_cull_grid(571, 821, seed=5, outlier_frac=0.05)
    dev = track.Device(0)
    lat = []
    for k in range(args.warmup + args.steps):
        cp, arrs = track.cull_params(offR, offA, corr, typ, islandThresh=20)
        t = time.perf_counter()
        dev.cull("st_cull_run", cp)
        if k >= args.warmup:
            lat.append(time.perf_counter() - t)
    n = offR.size
    line = {
        "metric": "cullst grid points/sec, 1 MI355X", "value": n * len(lat) / sum(lat), "unit": "grid points/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * sum(lat) / len(lat),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 (f64 plane fits)",
        "data": "synthetic", "config": {"workload": "cullst full-scene grid 821 x 571, 5% outliers, islands 20"},
        "latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)), "p99": 1e3 * float(np.percentile(lat, 99))},
        "note": "host grids in and out (PCIe included)",
    }
    dev.close()
    print(json.dumps(line))


def cpu_baseline(spec, par, dev, pid, budget_s=15.0):
    """Time the FP32 CPU restatement (single thread) on a strip of grid
    rows of the same resident scene, and use the first rows of that strip
    as the accuracy sample: the GPU rerun on the same strip with
    diagnostics is compared with the FP32 oracle (integer peaks, match
    types, near-ties) and the FP64 oracle (sub-pixel offsets,
    correlation)."""
    from oracle.pyoracle import Oracle
    from specklesource_amd import track
    from tests.helpers import compare_strack
    o = Oracle("f32")
    rows = 1
    total_pts = 0
    total_t = 0.0
    total_corr = 0
    i0 = par.nA // 3
    acc = None
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
        if acc is None and rows >= 2:
            spid = dev.upload_pair(img1, img2)
            g = dev.speckle_track(spid, sub, diag=True)
            dev.release_pair(spid)
            r64 = Oracle("f64").speckle_track(sub, img1, img2, diag=True)
            rep, probs = compare_strack(sub, g, r, r64)
            acc = {"sample_points": sub.nR * sub.nA,
                   "rms_offset_px_vs_fp64": [rep.get("rms_r_vs_f64"), rep.get("rms_a_vs_f64")],
                   "integer_peak_agreement_vs_fp32": rep["peak_agree"], "near_ties": rep["near_ties"],
                   "type_agreement": rep["type_agree"], "unexplained_mismatches": len(probs),
                   "corr_relerr_median_vs_fp64": rep.get("corr_relerr_median"),
                   "corr_relerr_max_vs_fp64": rep.get("corr_relerr_max")}
        i0 += rows
        rows = min(4, rows * 2)
    base = {"value": total_corr / total_t, "unit": "patch correlations/s", "cores": 1, "kind": "port",
            "sample": "%d grid points (%d consecutive grid rows of the same scene), %.1f s single-thread"
                      % (total_pts, total_pts // par.nR, total_t)}
    return base, acc


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=1, help="divide the scene sides (testing only)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=["served", "fullscene", "amponly", "strackw", "cull", "batched"],
                    default="served")
    ap.add_argument("--no-roofline", action="store_true", help="served workload: skip the one-context roofline pass")
    ap.add_argument("--roofline-only", action="store_true",
                    help="served workload: only the one-context roofline pass at --precision (for profilers)")
    ap.add_argument("--batch", type=int, default=1, help="served workload: strack jobs per fused batch (1: off)")
    ap.add_argument("--precision", choices=["fp32", "mfma"], default="fp32")
    ap.add_argument("--pairs", type=int, default=64)
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--windows", default="", help="batched workload: only these window sizes (e.g. 16,32)")
    ap.add_argument("--chip", type=int, default=128, help="strackw chip size (128 or 256)")
    ap.add_argument("--coherence", type=float, default=None,
                    help="full scene with this uniform coherence instead of 0.3-0.9 (e.g. 0.05: every point "
                         "falls back to amplitude matching)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a 1 GiB device copy (for scaling HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(each GPU would track its own share of the grid rows); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.workload == "strackw":
        bench_strackw(args)
        return
    if args.workload == "cull":
        bench_cull(args)
        return
    if args.workload == "batched":
        bench_batched(args)
        return
    if args.workload == "served":
        if args.roofline_only:
            from specklesource_amd import abi, configs
            pairs = configs.cfg3_pairs(n_pairs=args.pairs, scale=args.scale)
            prec = abi.ST_PREC_MFMA if args.precision == "mfma" else abi.ST_PREC_FP32
            print(json.dumps({"roofline": served_roofline(pairs, prec)}))
        else:
            bench_served(args)
        return

    from specklesource_amd import abi, configs, track
    amponly = args.workload == "amponly"
    spec, par = configs.cfg1(scale=args.scale) if amponly else configs.cfg2(scale=args.scale)
    if args.coherence is not None:
        spec = abi.SynthSpec.from_buffer_copy(spec)
        spec.gmin = spec.gmax = args.coherence
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
        for i in range(len(st.kernelNames)):
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
    if amponly:
        # amplitude-only job: the whole amplitude stage (k_amp_* and the
        # transforms) against the algorithmic flops of its tries
        kname = "amplitude stage (k_amp_*)"
        launch_ms = (stages[2] + stages[3]) / args.steps
        per_launch = flops_amp_patch(par.wAa) * st.nCorr[1] + flops_amp_patch(2 * par.wAa) * st.nCorr[2]
    elif "k_cpx_fused" in kern and kern["k_cpx_fused"][1] > 0:
        # the fused kernel and the refinement kernel that follows it
        kname = "k_cpx_fused + k_cpx_refine"
        launch_ms = (kern["k_cpx_fused"][0] + kern.get("k_cpx_refine", (0.0, 0))[0]) / kern["k_cpx_fused"][1]
        per_launch = ((flops_complex_patch(par.wR) + flops_window_sums(par.wR)) * ncpx * args.steps /
                      kern["k_cpx_fused"][1])
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
        "config": {"workload": ("amplitude-only single pair (-noComplex), 32x32 windows, 64x64 retries, 16 px grid"
                                if amponly else
                                "full-scene single pair, strack complex + amplitude, 64x64 windows, 24 px grid"),
                   "scene": [spec.nr, spec.na], "grid": [par.nR, par.nA], "window": par.wR,
                   "coherence": [spec.gmin, spec.gmax]},
        "latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)), "p99": 1e3 * float(np.percentile(lat, 99))},
        "work": {"patch_correlations_per_job": ncorr // args.steps, "complex": int(ncpx),
                 "amp_small": int(st.nCorr[1]), "amp_large": int(st.nCorr[2]),
                 "matches": {"complex": res.nComplex, "amp": res.nAmp, "ampL": res.nAmpL, "fail": res.nFail}},
        "stage_ms": {"geometry": stages[0] / args.steps, "complex": stages[1] / args.steps,
                     "amp_small": stages[2] / args.steps, "amp_large": stages[3] / args.steps},
        "kernels_ms_per_step": {k: v[0] / args.steps for k, v in kern.items()},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None if amponly else measured_traffic(),
                     "kernel": kname, "launch_ms": launch_ms, "flops_per_launch": per_launch,
                     "algorithmic_bytes_per_launch": (2 * 8 * par.wAa * par.wAa * (st.nCorr[1] + 4 * st.nCorr[2])
                                                      if amponly else 2 * 8 * par.wR * par.wR * ncpx),
                     "dominant_kernel": dom[0]},
    }
    if not args.no_cpu:
        line["cpu_baseline"], line["accuracy"] = cpu_baseline(spec, par, dev, pid)
    dev.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
