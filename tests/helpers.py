"""Shared helpers for the oracle pinning tests and the GPU parity tests."""
import ctypes as C

import numpy as np

from specklesource_amd import abi, configs, track

# Relative gap between the best and the runner-up correlation value below
# which two FP32 implementations may legitimately pick different peaks
# (FP32 FFT rounding on surfaces of <= 512^2 points is ~1e-6 relative).
TIE = 2e-5
# Relative distance to an acceptance threshold that FP32 rounding can cross.
THRESH_TOL = 1e-4
# Relative tolerance on the 1-D Gaussian widths (gaussPeak1D's
# Levenberg-Marquardt fit, five float iterations over exp()) and on their
# ratio, the .azd product: the GPU's exp and 3x3 solve round differently
# from the CPU's in the last bits, and five iterations amplify that.
LM_RTOL = 2e-3
# Relative tolerance on the SNR diagnostic (peak over the mean |C| of the
# surface: an FP32 sum over up to 512^2 values against the FP64 oracle's).
SNR_RTOL = 1e-3


def periodic_pair(w, nr, na, dr, da, seed=0, band=True):
    """Image 1 tiles one wA x wR pattern (w is wA = wR or a pair (wA, wR));
    image 2 is image 1 shifted by the integer offset (dr, da), so every
    wA x wR window of image 2 is a circular shift of the matching window of
    image 1 (a closed-form case: the correlation peak is exactly at the
    shift)."""
    wa, wr = (w, w) if np.isscalar(w) else w
    rng = np.random.default_rng(seed)
    pat = (rng.standard_normal((wa, wr)) + 1j * rng.standard_normal((wa, wr))) / np.sqrt(2)
    if band:
        # keep the pattern band-limited around a Doppler centroid
        F = np.fft.fft2(pat)
        fa = np.fft.fftfreq(wa)[:, None]
        fr = np.fft.fftfreq(wr)[None, :]
        keep = (np.abs(((fa - 0.12 + 0.5) % 1.0) - 0.5) < 0.4) & (np.abs(fr) < 0.43)
        pat = np.fft.ifft2(F * keep)
    reps = (na // wa + 2, nr // wr + 2)
    big = np.tile(pat, reps)
    img1 = big[:na, :nr].astype(np.complex64)
    img2 = np.roll(np.roll(big, da, axis=0), dr, axis=1)[:na, :nr].astype(np.complex64)
    return np.ascontiguousarray(img1), np.ascontiguousarray(img2)


def small_params(w, nr, na, step=None, **kw):
    step = step or w
    pad = 2 * w + 8
    nR = max(1, (nr - 2 * pad) // step + 1)
    nA = max(1, (na - 2 * pad) // step + 1)
    base = dict(rStart=pad, aStart=pad, deltaR=step, deltaA=step, nR=nR, nA=nA, wR=w, wA=w, wRa=w,
                wAa=w, rShiftPoly=[0, 0, 0], aShiftPoly=[0, 0, 0])
    base.update(kw)
    return track.default_params(**base)


def truth_grid(spec, par):
    """True offset at each grid point (patch centre)."""
    tr = np.zeros((par.nA, par.nR))
    ta = np.zeros((par.nA, par.nR))
    for i in range(par.nA):
        for j in range(par.nR):
            tr[i, j], ta[i, j] = track.synth_offset(spec, par.rStart + j * par.deltaR,
                                                    par.aStart + i * par.deltaA)
    return tr, ta


def compare_strack(par, g, o32, o64, tie=TIE, thresh_tol=THRESH_TOL, fine_tie=1e-3, lm_rtol=LM_RTOL):
    """Parity report of a GPU run g against the FP32-faithful (o32) and FP64
    (o64) oracle runs (all with diagnostics).  Returns (report, problems):
    problems lists points whose disagreement is not explained by a near-tie
    (relative top-two gap below tie, fine-surface gap below fine_tie) or a
    value within thresh_tol (relative) of an acceptance threshold."""
    rep = {}
    probs = []
    n = par.nR * par.nA
    d32, dg = o32.diag, g.diag
    # integer-peak agreement per attempt
    tot = agree = ties = 0
    tie_pts = set()
    for t in range(3):
        ran = (d32["peakI"][t] > -2000) & (dg["peakI"][t] > -2000)
        same = ran & (d32["peakI"][t] == dg["peakI"][t]) & (d32["peakJ"][t] == dg["peakJ"][t])
        diff = ran & ~same
        near = diff & ((d32["peakGap"][t] < tie) | (dg["peakGap"][t] < tie))
        tot += int(ran.sum())
        agree += int(same.sum())
        ties += int(near.sum())
        for idx in zip(*np.nonzero(near)):
            tie_pts.add(idx)
        for idx in zip(*np.nonzero(diff & ~near)):
            probs.append(("peak", t, idx, int(d32["peakI"][t][idx]), int(d32["peakJ"][t][idx]),
                          int(dg["peakI"][t][idx]), int(dg["peakJ"][t][idx]),
                          float(d32["peakGap"][t][idx])))
    rep["attempts"] = tot
    rep["peak_agree"] = agree / max(tot, 1)
    rep["near_ties"] = ties
    # match types
    tmis = g.type != o32.type
    expl = np.zeros_like(tmis)
    for idx in zip(*np.nonzero(tmis)):
        if idx in tie_pts:
            expl[idx] = True
            continue
        c = [d32["cmax"][t][idx] for t in range(3)]
        for t, thr in ((0, 0.185), (1, 0.07), (2, 0.028)):
            if abs(c[t] - thr) < thresh_tol * thr:
                expl[idx] = True
        # fine peak moved by one quantum near the acceptance limit
        if any(d32["fineGap"][t][idx] < fine_tie for t in range(3)):
            expl[idx] = True
    for idx in zip(*np.nonzero(tmis & ~expl)):
        probs.append(("type", idx, int(o32.type[idx]), int(g.type[idx])))
    rep["type_agree"] = 1.0 - tmis.sum() / n
    rep["type_mismatch_explained"] = int((tmis & expl).sum())
    # Gaussian widths of every attempt whose peak agrees, and the .azd
    # product (aZDefocus, speckleTrack.c:380-387 complex, :537-545
    # amplitude) where every attempt that ran agrees
    all_same = ~tmis
    lm_err = 0.0
    for t in range(3):
        ran = (d32["peakI"][t] > -2000) & (dg["peakI"][t] > -2000)
        same = ran & (d32["peakI"][t] == dg["peakI"][t]) & (d32["peakJ"][t] == dg["peakJ"][t])
        all_same &= ~(((d32["peakI"][t] > -2000) | (dg["peakI"][t] > -2000)) & ~same)
        for nm in ("sigRg", "sigAz"):
            a, b = d32[nm][t][same].astype(np.float64), dg[nm][t][same].astype(np.float64)
            e = np.abs(a - b) / np.maximum(np.abs(a), 1e-6)
            lm_err = max(lm_err, float(e.max()) if e.size else 0.0)
            for k in np.nonzero(e > lm_rtol)[0]:
                probs.append((nm, t, float(a[k]), float(b[k])))
    za, zb = o32.azd[all_same].astype(np.float64), g.azd[all_same].astype(np.float64)
    ez = np.abs(za - zb) / np.maximum(np.abs(za), 1e-6)
    # SNR against the FP64 oracle wherever the attempt's peak agrees
    d64 = o64.diag
    snr_err = []
    for t in range(3):
        same = (dg["peakI"][t] > -1000) & (d64["peakI"][t] == dg["peakI"][t]) & (d64["peakJ"][t] == dg["peakJ"][t])
        a, b = dg["snr"][t][same].astype(np.float64), d64["snr"][t][same].astype(np.float64)
        e = np.abs(a - b) / np.maximum(np.abs(b), 1e-12)
        snr_err += list(e)
        for k in np.nonzero(e > SNR_RTOL)[0]:
            probs.append(("snr", t, float(b[k]), float(a[k])))
    rep["snr_relerr_median"] = float(np.median(snr_err)) if snr_err else None
    rep["snr_relerr_max"] = float(np.max(snr_err)) if snr_err else None
    rep["lm_sigma_relerr_max"] = lm_err
    rep["azd_compared"] = int(all_same.sum())
    rep["azd_relerr_max"] = float(ez.max()) if ez.size else 0.0
    for k in np.nonzero(ez > lm_rtol)[0]:
        probs.append(("azd", float(za[k]), float(zb[k])))
    # sub-pixel offsets against the FP64 oracle
    both = (g.type > 0) & (o64.type > 0) & (g.type == o64.type)
    if both.any():
        er = (g.offR - o64.offR)[both]
        ea = (g.offA - o64.offA)[both]
        rep["rms_r_vs_f64"] = float(np.sqrt(np.mean(er.astype(np.float64) ** 2)))
        rep["rms_a_vs_f64"] = float(np.sqrt(np.mean(ea.astype(np.float64) ** 2)))
        cr = np.abs(g.corr[both].astype(np.float64) - o64.corr[both]) / np.maximum(np.abs(o64.corr[both]), 1e-12)
        rep["corr_relerr_median"] = float(np.median(cr))
        rep["corr_relerr_max"] = float(np.max(cr))
        same32 = both & (g.type == o32.type)
        rep["offset_equal_f32"] = float(np.mean((g.offR == o32.offR)[same32] & (g.offA == o32.offA)[same32]))
    rep["counts_gpu"] = (g.nComplex, g.nAmp, g.nAmpL, g.nFail)
    rep["counts_f32"] = (o32.nComplex, o32.nAmp, o32.nAmpL, o32.nFail)
    return rep, probs


# the cull workload generator and the CullParams builder live in the package
from specklesource_amd.configs import make_cull_grid  # noqa: E402,F401
from specklesource_amd.track import cull_params  # noqa: E402,F401


def fringe_scene(n, w, nrl, nal, kfa, kfr, seed, noise=0.2, step=None, second=None, chirp=False):
    """Synthetic pair whose image 1 carries the fringe of an interferogram
    with nrl x nal looks, and that interferogram ((na_i, 2 nr_i) float32
    pairs, amplitude 100).  The fringe has kfa / kfr cycles per ps
    interferogram samples, ps = max(w / nal, w / nrl, 8) (the patch of the
    correction), plus complex noise of relative amplitude `noise`.
    second = (kfa2, kfr2, amp2) adds a second fringe over the right half of
    the interferogram; chirp=True replaces the fringe by a quadratic-phase
    (flat-spectrum) pattern over the lower half."""
    ps = max(w // nal, w // nrl, 8)
    s = configs._spec(n, n, seed, dr0=0.6, da0=-0.4, gmin=0.8, gmax=0.9)
    a, b = track.synth_host(s)
    na_i, nr_i = n // nal + 2, n // nrl + 2
    ia = np.arange(na_i, dtype=np.float64)[:, None]
    ir = np.arange(nr_i, dtype=np.float64)[None, :]
    z = np.exp(2j * np.pi * (kfa * ia + kfr * ir) / ps)
    if second is not None:
        k2a, k2r, amp2 = second
        half = (ir >= nr_i // 2) * np.ones_like(ia)
        z = z + half * amp2 * np.exp(2j * np.pi * (k2a * ia + k2r * ir) / ps + 0.7j)
    if chirp:
        q = np.exp(1j * (np.pi * ((ia % ps) ** 2 + (ir % ps) ** 2) / ps + 0.3))
        low = (ia >= na_i // 2) * np.ones_like(ir)
        z = np.where(low > 0, q, z)
    rng = np.random.default_rng(seed)
    z = z + noise * (rng.standard_normal(z.shape) + 1j * rng.standard_normal(z.shape)) / np.sqrt(2)
    intf = np.empty((na_i, 2 * nr_i), np.float32)
    intf[:, 0::2] = 100 * z.real
    intf[:, 1::2] = 100 * z.imag
    # image 1 carries the interferogram's phase (image 1 x conj(image 2))
    aa = np.arange(n, dtype=np.float64)[:, None] / nal
    rr = np.arange(n, dtype=np.float64)[None, :] / nrl
    ph = np.angle(np.exp(2j * np.pi * (kfa * aa + kfr * rr) / ps))
    a = (a * np.exp(1j * ph)).astype(np.complex64)
    par = small_params(w, n, n, step=step or w // 2, rShiftPoly=[0.6, 0, 0], aShiftPoly=[-0.4, 0, 0])
    par.intf = abi.ptr(intf, C.c_float)
    par.intNa, par.intNr, par.intNrl, par.intNal = na_i, nr_i, nrl, nal
    return par, a, b, intf, ps
