"""Synthetic workloads of the five benchmark configurations.

Each function returns (spec, params) -- an abi.SynthSpec describing the
scene and the abi.TrackParams of the tracking run -- at full size or
scaled down for tests (`scale` divides the scene sides; windows, grid
steps and offsets keep their full-size values so the per-patch work is
the same).

  cfg1  strack -noComplex on 2048 x 4096, w=32, grid 16, affine offsets
  cfg2  strack complex+amplitude on 20000 x 14000, w=64, grid 24,
        glacier ridge (12 px az / 4 px rg peak), coherence 0.3-0.9,
        10% zero-filled
  cfg3  64 pairs 1024-8192 px, w in {16,32,64,128}, grid w/2, affine
        offsets within +/-20 px
  cfg4  strackw on 8192 x 8192, coherence 0.15-0.4, shear to +/-120 px,
        5x5 boxcar (navg=2), chip 128 or 256 with edge pad 123
"""
import numpy as np

from . import abi
from .track import default_params


def _spec(nr, na, seed, **kw):
    s = abi.SynthSpec()
    s.nr, s.na, s.seed = int(nr), int(na), int(seed)
    s.gmin, s.gmax = 0.7, 0.7
    s.sigmaLn, s.corrLen = 0.5, 64.0
    s.bwR, s.bwA, s.fdc = 0.85, 0.8, 0.12
    s.maskR0 = s.maskR1 = s.maskA0 = s.maskA1 = 0
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def grid_for(nr, na, w, step, pad):
    """Grid start/count covering the scene with `pad` px margins."""
    r0 = pad + w
    a0 = pad + w
    nR = max(1, (nr - 2 * r0) // step + 1)
    nA = max(1, (na - 2 * a0) // step + 1)
    return r0, a0, nR, nA


def affine_poly(spec):
    """Initial-offset polynomials p0 + p1*r + p2*a equal to the affine
    part of the synthetic field (the reference's initialoffsetfile)."""
    return [spec.dr0, spec.drR, spec.drA], [spec.da0, spec.daR, spec.daA]


def cfg1(scale=1, seed=1):
    nr, na = 2048 // scale, 4096 // scale
    s = _spec(nr, na, seed, dr0=1.0, drR=2.0 / 2048, da0=0.5, daA=2.0 / 4096, gmin=0.7, gmax=0.7)
    w, step = 32, 16
    r0, a0, nR, nA = grid_for(nr, na, 2 * w, step, 16)
    rp, ap = [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]
    p = default_params(rStart=r0, aStart=a0, deltaR=step, deltaA=step, nR=nR, nA=nA,
                       wR=w, wA=w, wRa=w, wAa=w, noComplex=1, rShiftPoly=rp, aShiftPoly=ap)
    return s, p


def cfg2(scale=1, seed=2, masked=True):
    nr, na = 20000 // scale, 14000 // scale
    kw = dict(dr0=0.3, drR=0.7 / 20000, da0=0.2, daA=0.8 / 14000,
              ridgeR=4.0, ridgeA=12.0, ridgeC=0.55 * nr, ridgeW=0.08 * nr,
              gmin=0.3, gmax=0.9, gLr=nr / 1.5, gLa=na / 1.25)
    if masked:
        kw.update(maskR0=int(0.75 * nr), maskR1=int(0.95 * nr), maskA0=int(0.25 * na),
                  maskA1=int(0.75 * na))
    s = _spec(nr, na, seed, **kw)
    w, step = 64, 24
    r0, a0, nR, nA = grid_for(nr, na, 2 * w, step, 24)
    rp, ap = affine_poly(s)
    p = default_params(rStart=r0, aStart=a0, deltaR=step, deltaA=step, nR=nR, nA=nA,
                       wR=w, wA=w, wRa=w, wAa=w, rShiftPoly=rp, aShiftPoly=ap)
    return s, p


def cfg3_pairs(n_pairs=64, scale=1, seed=3):
    """List of (spec, params) for the batched uneven workload."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_pairs):
        nr = int(rng.integers(1024, 8193)) // scale
        na = int(rng.integers(1024, 8193)) // scale
        w = int(rng.choice([16, 32, 64, 128]))
        step = w // 2
        c = rng.uniform(-20, 20, size=2)
        g = rng.uniform(-2e-3, 2e-3, size=4)
        s = _spec(nr, na, seed * 1000 + k, dr0=c[0], drR=g[0], drA=g[1], da0=c[1], daR=g[2],
                  daA=g[3], gmin=0.5, gmax=0.8, gLr=nr / 1.3, gLa=na / 1.7)
        pad = 2 * w + 40
        r0, a0, nR, nA = grid_for(nr, na, w, step, pad)
        rp, ap = affine_poly(s)
        p = default_params(rStart=r0, aStart=a0, deltaR=step, deltaA=step, nR=nR, nA=nA,
                           wR=w, wA=w, wRa=w, wAa=w, rShiftPoly=rp, aShiftPoly=ap)
        out.append((s, p))
    return out


def cfg4(chip=128, scale=1, seed=4):
    nr = na = 8192 // scale
    s = _spec(nr, na, seed, shearA=120.0 * min(1.0, 1.0), dr0=0.5, gmin=0.15, gmax=0.4,
              gLr=nr / 1.1, gLa=na / 1.3)
    edge = 123
    wRa = chip + 2 * (edge + 5)
    step = chip // 2
    r0, a0, nR, nA = grid_for(nr, na, wRa, step, 150)
    p = default_params(rStart=r0, aStart=a0, deltaR=step, deltaA=step, nR=nR, nA=nA,
                       wR=wRa, wA=wRa, wRa=wRa, wAa=wRa, edgePadR=edge, edgePadA=edge,
                       navgR=2, navgA=2, noComplex=1)
    return s, p


def initial_grid(spec, par, factor=8, err_sigma=2.5, err_clip=5.0, seed=44):
    """Initial-offset grid at 1/factor of the output resolution with
    clipped Gaussian error (cfg4's supplied estimate)."""
    from .track import synth_offset
    rng = np.random.default_rng(seed)
    nR0 = max(2, par.nR // factor + 2)
    nA0 = max(2, par.nA // factor + 2)
    dR = par.deltaR * factor
    dA = par.deltaA * factor
    r0 = par.rStart - par.wRa // 2
    a0 = par.aStart - par.wAa // 2
    gr = np.zeros((nA0, nR0), np.float32)
    ga = np.zeros((nA0, nR0), np.float32)
    for i in range(nA0):
        for j in range(nR0):
            dr, da = synth_offset(spec, r0 + j * dR + par.wRa / 2, a0 + i * dA + par.wAa / 2)
            e = np.clip(rng.normal(0, err_sigma, 2), -err_clip, err_clip)
            gr[i, j] = dr + e[0]
            ga[i, j] = da + e[1]
    return gr, ga, (r0, a0, dR, dA)


def make_cull_grid(nA, nR, seed=0, outlier_frac=0.05, hole_frac=0.08):
    """Offset grids like strack output: smooth field + noise, a fraction of
    uniform +/-30 px outliers, random holes (-2e9) and islands."""
    rng = np.random.default_rng(seed)
    ii, jj = np.mgrid[0:nA, 0:nR]
    offR = (1.5 + 0.004 * jj + 0.5 * np.sin(ii / 17.0) + 0.05 * rng.standard_normal((nA, nR))).astype(np.float32)
    offA = (-0.8 + 0.003 * ii + 0.4 * np.cos(jj / 23.0) + 0.05 * rng.standard_normal((nA, nR))).astype(np.float32)
    out = rng.random((nA, nR)) < outlier_frac
    offR[out] += rng.uniform(-30, 30, out.sum()).astype(np.float32)
    offA[out] += rng.uniform(-30, 30, out.sum()).astype(np.float32)
    holes = rng.random((nA, nR)) < hole_frac
    # a block of missing data with a few small islands inside
    holes[nA // 3: nA // 3 + nA // 5, nR // 4: nR // 4 + nR // 3] = True
    for k in range(4):
        a0 = nA // 3 + 2 + 4 * k
        r0 = nR // 4 + 3 + 7 * k
        holes[a0:a0 + 2 + k, r0:r0 + 3] = False
    offR[holes] = -2e9
    offA[holes] = -2e9
    typ = rng.integers(1, 4, (nA, nR)).astype(np.uint8)
    typ[holes] = 0
    corr = rng.uniform(0.0, 1.0, (nA, nR)).astype(np.float32)
    return offR, offA, corr, typ


def inject_outliers(offR, offA, frac=0.05, seed=0, amp=30.0):
    """Copies of the offset grids with `frac` of the valid points moved by
    uniform +/-amp px in both components (the served stream's injected
    outliers)."""
    rng = np.random.default_rng(seed)
    r, a = offR.copy(), offA.copy()
    pick = (rng.random(r.shape) < frac) & (r > -1e9) & (a > -1e9)
    r[pick] += rng.uniform(-amp, amp, int(pick.sum())).astype(np.float32)
    a[pick] += rng.uniform(-amp, amp, int(pick.sum())).astype(np.float32)
    return r, a


def cfg5_stream(n_jobs=40, scale=1, seed=5, rate=20.0):
    """The served stream: Poisson arrivals at `rate` jobs/s of tracking jobs
    drawn from the other four configurations by size -- amplitude-only
    scenes (cfg1), complex + amplitude scenes up to 20000 px with 64 px
    windows (cfg2), uneven pairs with windows 16-128 (cfg3) and strackw
    scenes with chips 128 / 256 (cfg4) -- each with a cullst pass (box 5 or
    9, islands below 20).  Returns a list of dicts with keys kind ('strack'
    / 'strackw'), spec, par, arrival (s), cull (keyword overrides of the
    cull parameters), src (the configuration it was drawn from)."""
    rng = np.random.default_rng(seed)
    jobs, t = [], 0.0
    for k in range(n_jobs):
        t += float(rng.exponential(1.0 / rate))
        src = int(rng.integers(1, 5))
        sub = seed * 100 + k
        if src == 1:
            s, p = cfg1(scale=scale, seed=sub)
            kind = "strack"
        elif src == 2:
            side = int(rng.integers(4096, 20001))
            nr, na = side // scale, int(side * 0.7) // scale
            s = _spec(nr, na, sub, dr0=0.3, da0=0.2, ridgeR=4.0, ridgeA=12.0, ridgeC=0.55 * nr,
                      ridgeW=0.08 * nr, gmin=0.3, gmax=0.9, gLr=nr / 1.5, gLa=na / 1.25)
            r0, a0, nR, nA = grid_for(nr, na, 128, 24, 24)
            rp, ap = affine_poly(s)
            p = default_params(rStart=r0, aStart=a0, deltaR=24, deltaA=24, nR=nR, nA=nA, wR=64, wA=64,
                               wRa=64, wAa=64, rShiftPoly=rp, aShiftPoly=ap)
            kind = "strack"
        elif src == 3:
            (s, p), = cfg3_pairs(n_pairs=1, scale=scale, seed=sub)
            kind = "strack"
        else:
            s, p = cfg4(chip=int(rng.choice([128, 256])), scale=scale, seed=sub)
            kind = "strackw"
        box = 5 if k % 2 else 9
        jobs.append(dict(kind=kind, spec=s, par=p, arrival=t, src=src,
                         cull=dict(bR=box, bA=box, islandThresh=20)))
    return jobs
