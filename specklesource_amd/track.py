"""Host-side mirror of the reference tracking interface.

The reference drives its hot path through three calls on fat structs:
``speckleTrack(TrackParams*)`` (Strack/strack.h:170),
``corrTrackFast(TrackParams*)`` (Strackw/strackw.h:1) and the cull
functions on ``CullParams*`` (Cullst/cullst.h:60-67).  This module keeps
those names and argument meanings for Python callers (tests, bench,
serving) and runs them on the MI355X library through its C ABI.  The
flag defaults are the reference's (Strack/strack.c:143-154,
Cullst/cullst.c:140-150).
"""
import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libspeckle_amd.so")

_lib = None


def load_library(path=LIB_PATH):
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            "libspeckle_amd.so is not built (%s); run __graft_entry__.build()" % path)
    lib = C.CDLL(path)
    vp = C.c_void_p
    lib.st_ctx_create.argtypes = [C.c_int32, C.POINTER(vp)]
    lib.st_ctx_destroy.argtypes = [vp]
    lib.st_last_error.argtypes = [vp]
    lib.st_last_error.restype = C.c_char_p
    lib.st_pair_upload.argtypes = [vp, vp, C.c_int32, C.c_int32, vp, C.c_int32,
                                   C.c_int32, C.c_int32, C.POINTER(C.c_int32)]
    lib.st_pair_upload_raw.argtypes = [vp, vp, C.c_int32, C.c_int32, vp, C.c_int32, C.c_int32, C.c_int32,
                                       C.c_int32, C.POINTER(C.c_int32)]
    lib.st_pair_release.argtypes = [vp, C.c_int32]
    lib.st_pair_download.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, vp, vp]
    lib.st_speckle_track.argtypes = [vp, C.c_int32, C.POINTER(abi.TrackParams),
                                     C.POINTER(abi.TrackOut), C.POINTER(abi.TrackDiag)]
    lib.st_corr_track_fast.argtypes = lib.st_speckle_track.argtypes
    lib.st_speckle_track_batch.argtypes = [vp, C.c_int32, C.POINTER(C.c_int32), C.POINTER(abi.TrackParams),
                                           C.POINTER(abi.TrackOut)]
    for name in ("st_cull_st_data", "st_cull_stats", "st_cull_smooth",
                 "st_cull_islands", "st_cull_run", "st_cull_run_noalloc"):
        getattr(lib, name).argtypes = [vp, C.POINTER(abi.CullParams)]
    for name in ("st_cull_ls_data", "st_cull_ls_stats", "st_cull_ls_smooth",
                 "st_cull_ls_islands", "st_cull_ls_run"):
        getattr(lib, name).argtypes = [vp, C.POINTER(abi.CullLSParams)]
    lib.st_last_stats.argtypes = [vp, C.POINTER(abi.TrackStats)]
    lib.st_server_create.argtypes = [C.c_int32, C.c_int32, C.c_uint64, C.POINTER(vp)]
    lib.st_server_destroy.argtypes = [vp]
    lib.st_server_pair_upload.argtypes = [vp, vp, C.c_int32, C.c_int32, vp, C.c_int32, C.c_int32, C.c_int32,
                                          C.POINTER(C.c_int32)]
    lib.st_server_pair_release.argtypes = [vp, C.c_int32]
    lib.st_server_pair_synth.argtypes = [vp, C.POINTER(abi.SynthSpec), C.POINTER(C.c_int32)]
    lib.st_submit.argtypes = [vp, C.POINTER(abi.Job), C.POINTER(C.c_int64)]
    lib.st_poll.argtypes = [vp, C.c_int64, C.POINTER(abi.JobInfo)]
    lib.st_wait.argtypes = [vp, C.c_int64, C.POINTER(abi.JobInfo)]
    lib.st_cancel.argtypes = [vp, C.c_int64]
    lib.st_autotune.argtypes = [vp, C.POINTER(C.c_int32), C.c_int32, C.c_char_p]
    lib.st_tune_get.argtypes = [vp, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                C.POINTER(C.c_float)]
    lib.st_tune_load.argtypes = [vp, C.c_char_p]
    lib.st_server_set_batch.argtypes = [vp, C.c_int32]
    lib.st_job_release.argtypes = [vp, C.c_int64]
    lib.st_set_profiling.argtypes = [vp, C.c_int32]
    lib.st_synth_host.argtypes = [C.POINTER(abi.SynthSpec), vp, vp]
    lib.st_synth_device.argtypes = [vp, C.POINTER(abi.SynthSpec), C.POINTER(C.c_int32)]
    lib.st_synth_offset.argtypes = [C.POINTER(abi.SynthSpec), C.c_double, C.c_double,
                                    C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.st_synth_offset.restype = None
    _lib = lib
    return lib


def default_params(**kw):
    """TrackParams with the reference defaults (strack.c:143-154)."""
    p = abi.TrackParams()
    p.scaleFactor = 1
    p.hanningFlag = 1
    p.maxTries = 2
    p.azDefocusThresh = 1.8
    p.polyShift = 1
    p.precision = abi.ST_PREC_FP32
    for k, v in kw.items():
        if k in ("rShiftPoly", "aShiftPoly"):
            getattr(p, k)[:] = list(v)
        else:
            setattr(p, k, v)
    return p


@dataclass
class TrackResult:
    offR: np.ndarray
    offA: np.ndarray
    corr: np.ndarray
    azd: np.ndarray
    type: np.ndarray
    nComplex: int = 0
    nAmp: int = 0
    nAmpL: int = 0
    nFail: int = 0
    diag: dict = field(default_factory=dict)


def alloc_outputs(par, want_diag=False):
    shape = (par.nA, par.nR)
    res = TrackResult(offR=np.zeros(shape, np.float32), offA=np.zeros(shape, np.float32),
                      corr=np.zeros(shape, np.float32), azd=np.zeros(shape, np.float32),
                      type=np.zeros(shape, np.uint8))
    out = abi.TrackOut()
    out.offR = abi.ptr(res.offR, C.c_float)
    out.offA = abi.ptr(res.offA, C.c_float)
    out.corr = abi.ptr(res.corr, C.c_float)
    out.azd = abi.ptr(res.azd, C.c_float)
    out.type = abi.ptr(res.type, C.c_uint8)
    dg = None
    if want_diag:
        dg = abi.TrackDiag()
        d = {}
        for nm in ("r2", "a2", "azShift"):
            d[nm] = np.zeros(shape, np.int32)
            setattr(dg, nm, abi.ptr(d[nm], C.c_int32))
        for nm, dt, ct in (("peakI", np.int32, C.c_int32), ("peakJ", np.int32, C.c_int32),
                           ("fineI", np.int32, C.c_int32), ("fineJ", np.int32, C.c_int32),
                           ("cmax", np.float64, C.c_double), ("sigRg", np.float32, C.c_float),
                           ("sigAz", np.float32, C.c_float), ("peakGap", np.float32, C.c_float),
                           ("fineGap", np.float32, C.c_float), ("snr", np.float32, C.c_float)):
            arrs = [np.zeros(shape, dt) for _ in range(3)]
            d[nm] = arrs
            fld = getattr(dg, nm)
            for t in range(3):
                fld[t] = abi.ptr(arrs[t], ct)
        res.diag = d
    return res, out, dg


def finish(res, out):
    res.nComplex, res.nAmp, res.nAmpL, res.nFail = out.nComplex, out.nAmp, out.nAmpL, out.nFail
    return res


class Device:
    """One MI355X context (st_ctx) with resident SLC pairs."""

    def __init__(self, device=0):
        self.lib = load_library()
        self.ctx = C.c_void_p()
        rc = self.lib.st_ctx_create(device, C.byref(self.ctx))
        if rc != abi.ST_OK:
            raise RuntimeError("st_ctx_create failed (%d)" % rc)

    def _check(self, rc, what):
        if rc != abi.ST_OK:
            msg = self.lib.st_last_error(self.ctx)
            raise RuntimeError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    def close(self):
        if self.ctx:
            self.lib.st_ctx_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload_pair(self, img1, img2, sample_type=abi.ST_SAMPLE_C64):
        img1 = np.ascontiguousarray(img1)
        img2 = np.ascontiguousarray(img2)
        pid = C.c_int32()
        rc = self.lib.st_pair_upload(self.ctx, img1.ctypes.data, img1.shape[1], img1.shape[0],
                                     img2.ctypes.data, img2.shape[1], img2.shape[0],
                                     sample_type, C.byref(pid))
        self._check(rc, "st_pair_upload")
        return pid.value

    def upload_pair_raw(self, raw1, shape1, raw2, shape2, sample_type=abi.ST_SAMPLE_C64, byte_order=1):
        """Upload samples as the reference's files hold them (bytes objects or
        uint8 arrays; shapes (na, nr)): c64 big-endian when byte_order is
        MSB (1), ci16 always big-endian; the swap runs on the GPU."""
        b1 = np.frombuffer(raw1, np.uint8) if isinstance(raw1, (bytes, bytearray)) else np.ascontiguousarray(raw1)
        b2 = np.frombuffer(raw2, np.uint8) if isinstance(raw2, (bytes, bytearray)) else np.ascontiguousarray(raw2)
        es = 8 if sample_type == abi.ST_SAMPLE_C64 else 4
        if b1.nbytes != es * shape1[0] * shape1[1] or b2.nbytes != es * shape2[0] * shape2[1]:
            raise ValueError("raw buffer sizes do not match the shapes")
        pid = C.c_int32()
        rc = self.lib.st_pair_upload_raw(self.ctx, b1.ctypes.data, shape1[1], shape1[0], b2.ctypes.data, shape2[1],
                                         shape2[0], sample_type, byte_order, C.byref(pid))
        self._check(rc, "st_pair_upload_raw")
        return pid.value

    def synth_pair(self, spec):
        pid = C.c_int32()
        self._check(self.lib.st_synth_device(self.ctx, C.byref(spec), C.byref(pid)),
                    "st_synth_device")
        return pid.value

    def download_pair(self, pair, row0, nrows, nr):
        a = np.zeros((nrows, nr), np.complex64)
        b = np.zeros((nrows, nr), np.complex64)
        self._check(self.lib.st_pair_download(self.ctx, pair, row0, nrows, a.ctypes.data,
                                              b.ctypes.data), "st_pair_download")
        return a, b

    def release_pair(self, pair):
        self._check(self.lib.st_pair_release(self.ctx, pair), "st_pair_release")

    def speckle_track(self, pair, par, diag=False):
        res, out, dg = alloc_outputs(par, diag)
        rc = self.lib.st_speckle_track(self.ctx, pair, C.byref(par), C.byref(out),
                                       C.byref(dg) if dg is not None else None)
        self._check(rc, "st_speckle_track")
        return finish(res, out)

    def autotune(self, windows, cache_file=None):
        """st_autotune over the given windows; writes cache_file when given."""
        w = (C.c_int32 * len(windows))(*windows)
        self._check(self.lib.st_autotune(self.ctx, w, len(windows), cache_file.encode() if cache_file else None),
                    "st_autotune")

    def tune(self, window):
        """(precision, wg_per_cu, staged_batch, us_per_patch) of a window, or
        None without an entry."""
        p, g, b, u = C.c_int32(), C.c_int32(), C.c_int32(), C.c_float()
        if self.lib.st_tune_get(self.ctx, window, C.byref(p), C.byref(g), C.byref(b), C.byref(u)) != abi.ST_OK:
            return None
        return p.value, g.value, b.value, u.value

    def tune_load(self, cache_file):
        self._check(self.lib.st_tune_load(self.ctx, cache_file.encode()), "st_tune_load")

    def speckle_track_batch(self, pairs, pars):
        """Several strack calls as one batch (st_speckle_track_batch); a
        list of TrackResult in the order of pairs / pars."""
        n = len(pairs)
        res, outs = [], (abi.TrackOut * n)()
        for k, p in enumerate(pars):
            r, o, _ = alloc_outputs(p)
            res.append(r)
            outs[k] = o
        pa = (abi.TrackParams * n)(*pars)
        pi = (C.c_int32 * n)(*pairs)
        self._check(self.lib.st_speckle_track_batch(self.ctx, n, pi, pa, outs), "st_speckle_track_batch")
        return [finish(r, outs[k]) for k, r in enumerate(res)]

    def corr_track_fast(self, pair, par, diag=False):
        res, out, dg = alloc_outputs(par, diag)
        rc = self.lib.st_corr_track_fast(self.ctx, pair, C.byref(par), C.byref(out),
                                         C.byref(dg) if dg is not None else None)
        self._check(rc, "st_corr_track_fast")
        return finish(res, out)

    def last_stats(self):
        st = abi.TrackStats()
        self._check(self.lib.st_last_stats(self.ctx, C.byref(st)), "st_last_stats")
        return st

    def set_profiling(self, on):
        self._check(self.lib.st_set_profiling(self.ctx, 1 if on else 0), "st_set_profiling")

    def cull(self, name, cp):
        self._check(getattr(self.lib, name)(self.ctx, C.byref(cp)), name)
        return cp


class Server:
    """Continuous-batching tracking server on one GPU (st_server): jobs of
    any scene and window size are submitted at once and run concurrently
    on `slots` streams, admitted by the HBM budget."""

    def __init__(self, device=0, slots=4, hbm_budget=200 << 30):
        self.lib = load_library()
        self.srv = C.c_void_p()
        rc = self.lib.st_server_create(device, slots, hbm_budget, C.byref(self.srv))
        if rc != abi.ST_OK:
            raise RuntimeError("st_server_create failed (%d)" % rc)
        self._jobs = {}

    def close(self):
        if self.srv:
            self.lib.st_server_destroy(self.srv)
            self.srv = C.c_void_p()
            self._jobs.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload_pair(self, img1, img2, sample_type=abi.ST_SAMPLE_C64):
        img1 = np.ascontiguousarray(img1)
        img2 = np.ascontiguousarray(img2)
        pid = C.c_int32()
        rc = self.lib.st_server_pair_upload(self.srv, img1.ctypes.data, img1.shape[1], img1.shape[0],
                                            img2.ctypes.data, img2.shape[1], img2.shape[0], sample_type,
                                            C.byref(pid))
        if rc != abi.ST_OK:
            raise RuntimeError("st_server_pair_upload failed (%d)" % rc)
        return pid.value

    def synth_pair(self, spec):
        """Generate a synthetic pair directly in HBM as a resident pair."""
        pid = C.c_int32()
        rc = self.lib.st_server_pair_synth(self.srv, C.byref(spec), C.byref(pid))
        if rc != abi.ST_OK:
            raise RuntimeError("st_server_pair_synth failed (%d)" % rc)
        return pid.value

    def release_pair(self, pair):
        rc = self.lib.st_server_pair_release(self.srv, pair)
        if rc != abi.ST_OK:
            raise RuntimeError("st_server_pair_release failed (%d)" % rc)

    def submit(self, par, pair=None, images=None, kind=abi.ST_JOB_STRACK, sample_type=abi.ST_SAMPLE_C64,
               cull=None):
        """Queue a job on a resident pair or on host images (img1, img2);
        cull = abi.CullParams over the job's output arrays (see
        cull_for_result) runs cullst after tracking.  Returns the job id."""
        res, out, _ = alloc_outputs(par)
        job = abi.Job()
        job.kind = kind
        keep = [res, par]
        if pair is not None:
            job.pair = pair
        else:
            img1 = np.ascontiguousarray(images[0])
            img2 = np.ascontiguousarray(images[1])
            keep += [img1, img2]
            job.pair = -1
            job.img1, job.img2 = img1.ctypes.data, img2.ctypes.data
            job.nr1, job.na1 = img1.shape[1], img1.shape[0]
            job.nr2, job.na2 = img2.shape[1], img2.shape[0]
            job.sampleType = sample_type
        job.par = par
        job.out = out
        if cull is not None:
            cp, arrs = cull(res)
            keep += [cp, arrs]
            job.cull = C.pointer(cp)
        jid = C.c_int64()
        rc = self.lib.st_submit(self.srv, C.byref(job), C.byref(jid))
        if rc != abi.ST_OK:
            raise RuntimeError("st_submit failed (%d)" % rc)
        self._jobs[jid.value] = (res, out, job, keep)
        return jid.value

    def submit_cull(self, cp, arrs):
        """Queue a cullst job on host grids (abi.CullParams from
        cull_params, with its arrays); wait() returns them culled as
        `extra`."""
        job = abi.Job()
        job.kind = abi.ST_JOB_CULL
        job.pair = -1
        job.cull = C.pointer(cp)
        jid = C.c_int64()
        rc = self.lib.st_submit(self.srv, C.byref(job), C.byref(jid))
        if rc != abi.ST_OK:
            raise RuntimeError("st_submit failed (%d)" % rc)
        self._jobs[jid.value] = (None, None, job, [cp, arrs])
        return jid.value

    def poll(self, jid):
        info = abi.JobInfo()
        if self.lib.st_poll(self.srv, jid, C.byref(info)) != abi.ST_OK:
            raise RuntimeError("unknown job %d" % jid)
        return info

    def set_batch(self, max_jobs):
        """Ready strack jobs on the fused path taken as one batch (1: off)."""
        if self.lib.st_server_set_batch(self.srv, max_jobs) != abi.ST_OK:
            raise RuntimeError("st_server_set_batch failed")

    def cancel(self, jid):
        """Ask the server to cancel a job (st_cancel); False when the job
        had already ended."""
        return self.lib.st_cancel(self.srv, jid) == abi.ST_OK

    def wait(self, jid, check=True):
        """Block until the job ends and forget it on the server
        (st_job_release); returns (TrackResult, JobInfo, extra) where extra
        holds the cull arrays when the job culled.  With check, a job that
        did not finish raises."""
        info = abi.JobInfo()
        if self.lib.st_wait(self.srv, jid, C.byref(info)) != abi.ST_OK:
            raise RuntimeError("unknown job %d" % jid)
        if self.lib.st_job_release(self.srv, jid) != abi.ST_OK:
            raise RuntimeError("st_job_release failed for job %d" % jid)
        res, out, job, keep = self._jobs.pop(jid)
        if check and info.state != abi.ST_JOB_DONE:
            raise RuntimeError("job %d failed (%d): %s" % (jid, info.status, info.error.decode()))
        if res is not None:
            res.nComplex, res.nAmp, res.nAmpL, res.nFail = info.nComplex, info.nAmp, info.nAmpL, info.nFail
        extra = keep[-1] if job.cull else None
        return res, info, extra


def cull_for_result(**kw):
    """Cull callback for Server.submit: cullst runs in place on the job's
    own output grids (reference defaults, overridden by kw)."""
    def make(res):
        nA, nR = res.offR.shape
        arrs = dict(offRS=np.zeros((nA, nR), np.float32), offAS=np.zeros((nA, nR), np.float32),
                    sigmaR=np.zeros((nA, nR), np.float32), sigmaA=np.zeros((nA, nR), np.float32))
        cp = abi.CullParams()
        cp.nR, cp.nA = nR, nA
        cp.offR, cp.offA = abi.ptr(res.offR, C.c_float), abi.ptr(res.offA, C.c_float)
        cp.corr, cp.type = abi.ptr(res.corr, C.c_float), abi.ptr(res.type, C.c_uint8)
        for k, v in arrs.items():
            setattr(cp, k, abi.ptr(v, C.c_float))
        cp.bR = cp.bA = 9
        cp.nGood = 17
        cp.maxA, cp.maxR = 1.0, 0.75
        cp.sR, cp.sA = 2, 8
        cp.islandThresh = -1
        cp.singleMT = -1
        cp.corrThresh = 0.0
        for k, v in kw.items():
            setattr(cp, k, v)
        return cp, arrs
    return make


def synth_host(spec):
    """Generate a synthetic pair on the host (complex64 arrays na x nr)."""
    lib = load_library()
    a = np.zeros((spec.na, spec.nr), np.complex64)
    b = np.zeros((spec.na, spec.nr), np.complex64)
    rc = lib.st_synth_host(C.byref(spec), a.ctypes.data, b.ctypes.data)
    if rc != abi.ST_OK:
        raise RuntimeError("st_synth_host failed (%d)" % rc)
    return a, b


def synth_offset(spec, r, a):
    lib = load_library()
    dr, da = C.c_double(), C.c_double()
    lib.st_synth_offset(C.byref(spec), float(r), float(a), C.byref(dr), C.byref(da))
    return dr.value, da.value


def cull_params(offR, offA, corr, typ, mask=None, **kw):
    """abi.CullParams over copies of the input grids (reference defaults,
    Cullst/cullst.c:140-150); returns (params, arrays) keeping buffers alive."""
    nA, nR = offR.shape
    arrs = dict(offR=offR.copy(), offA=offA.copy(), corr=corr.copy(), type=typ.copy(),
                offRS=np.zeros_like(offR), offAS=np.zeros_like(offR), sigmaR=np.zeros_like(offR),
                sigmaA=np.zeros_like(offR))
    if mask is not None:
        arrs["mask"] = mask.copy()
    cp = abi.CullParams()
    cp.nR, cp.nA = nR, nA
    for k, v in arrs.items():
        setattr(cp, k, abi.ptr(v, C.c_uint8 if v.dtype == np.uint8 else C.c_float))
    cp.bR = cp.bA = 9
    cp.nGood = 17
    cp.maxA, cp.maxR = 1.0, 0.75
    cp.sR, cp.sA = 2, 8
    cp.islandThresh = -1
    cp.singleMT = -1
    cp.corrThresh = 0.0
    for k, v in kw.items():
        setattr(cp, k, v)
    return cp, arrs


def cull_ls_params(X, Y, typ, *, slow_vel_mask, no_data_to_use_mask, no_data, nover, deltaT, dx, dy, **kw):
    """abi.CullLSParams over copies of Landsat offset grids (the cullls
    defaults of Cullls/cullls.c:146-153 for the rest, overridden by kw);
    returns (params, arrays) keeping the buffers alive.  The type bits,
    no-data value and oversampling factor are those of the caller's Landsat
    matcher (they live in headers outside the reference tree)."""
    ny, nx = X.shape
    if Y.shape != X.shape or typ.shape != X.shape:
        raise ValueError("X, Y and type must have the same shape")
    arrs = dict(X=X.astype(np.float32).copy(), Y=Y.astype(np.float32).copy(), type=typ.astype(np.uint8).copy(),
                XS=np.zeros((ny, nx), np.float32), YS=np.zeros((ny, nx), np.float32),
                sigmaX=np.zeros((ny, nx), np.float32), sigmaY=np.zeros((ny, nx), np.float32))
    cp = abi.CullLSParams()
    cp.nx, cp.ny = nx, ny
    for k, v in arrs.items():
        setattr(cp, k, abi.ptr(v, C.c_uint8 if v.dtype == np.uint8 else C.c_float))
    cp.bR = cp.bA = 9
    cp.nGood = 17
    cp.maxX, cp.maxY = 0.75, 1.0
    cp.sX, cp.sY = 3, 3
    cp.islandThresh = -1
    cp.deltaT, cp.dx, cp.dy = deltaT, dx, dy
    cp.slowVelMask, cp.noDataToUseMask = slow_vel_mask, no_data_to_use_mask
    cp.noData, cp.nover = no_data, nover
    cp.nMatch = -1
    for k, v in kw.items():
        setattr(cp, k, v)
    return cp, arrs
