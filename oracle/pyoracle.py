"""ORACLE loader -- test infrastructure, not the product.

Loads the CPU restatement of the reference hot path (oracle/*.c, built by
oracle/Makefile into oracle/_build/) and exposes it with the same call
shapes as specklesource_amd.track.Device.  Only tests/, the smoke check
and bench.py's cpu_baseline leg use it, as the checker or the CPU
baseline.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from specklesource_amd import abi
from specklesource_amd.track import alloc_outputs, finish

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(_HERE, "_build")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_libs = {}


def _load(prec):
    if prec in _libs:
        return _libs[prec]
    path = os.path.join(_BUILD, "liboracle_%s.so" % prec)
    if not os.path.exists(path):
        build()
    lib = C.CDLL(path)
    vp = C.c_void_p
    sig = [C.POINTER(abi.TrackParams), vp, C.c_int, C.c_int, vp, C.c_int, C.c_int, C.c_int,
           C.POINTER(abi.TrackOut), C.POINTER(abi.TrackDiag)]
    lib.orc_speckle_track.argtypes = sig
    lib.orc_corr_track_fast.argtypes = sig
    for nm in ("orc_cull_st_data", "orc_cull_stats", "orc_cull_smooth", "orc_cull_islands",
               "orc_cull_run"):
        getattr(lib, nm).argtypes = [C.POINTER(abi.CullParams)]
    for nm in ("orc_cull_ls_data", "orc_cull_ls_stats", "orc_cull_ls_smooth", "orc_cull_ls_islands",
               "orc_cull_ls_run"):
        getattr(lib, nm).argtypes = [C.POINTER(abi.CullLSParams)]
    lib.orc_fft1d.argtypes = [vp, C.c_int, C.c_int]
    lib.orc_fft2d.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    _libs[prec] = lib
    return lib


class Oracle:
    """prec = 'f32' (FP32-faithful restatement) or 'f64' (FP64 yardstick)."""

    def __init__(self, prec="f32"):
        self.prec = prec
        self.lib = _load(prec)
        self.dtype = np.complex64 if prec == "f32" else np.complex128

    def fft(self, x, sign=-1):
        """1-D or 2-D unnormalised FFT, FFTW 2 sign convention."""
        y = np.ascontiguousarray(x, dtype=self.dtype).copy()
        if y.ndim == 1:
            self.lib.orc_fft1d(y.ctypes.data, y.shape[0], sign)
        else:
            self.lib.orc_fft2d(y.ctypes.data, y.shape[0], y.shape[1], sign)
        return y

    def _track(self, fn, par, img1, img2, sample_type, diag):
        img1 = np.ascontiguousarray(img1)
        img2 = np.ascontiguousarray(img2)
        res, out, dg = alloc_outputs(par, diag)
        rc = fn(C.byref(par), img1.ctypes.data, img1.shape[1], img1.shape[0], img2.ctypes.data,
                img2.shape[1], img2.shape[0], sample_type, C.byref(out),
                C.byref(dg) if dg is not None else None)
        if rc != abi.ST_OK:
            raise RuntimeError("oracle tracking failed (%d)" % rc)
        return finish(res, out)

    def speckle_track(self, par, img1, img2, sample_type=abi.ST_SAMPLE_C64, diag=False):
        return self._track(self.lib.orc_speckle_track, par, img1, img2, sample_type, diag)

    def corr_track_fast(self, par, img1, img2, sample_type=abi.ST_SAMPLE_C64, diag=False):
        return self._track(self.lib.orc_corr_track_fast, par, img1, img2, sample_type, diag)

    def cull(self, name, cp):
        rc = getattr(self.lib, "orc_" + name[3:])(C.byref(cp))
        if rc != abi.ST_OK:
            raise RuntimeError("oracle %s failed (%d)" % (name, rc))
        return cp
