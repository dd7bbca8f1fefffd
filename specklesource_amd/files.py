"""Offset and cull products on disk (include/speckle_io.h), through the C
ABI: the files strack / strackw write after tracking (.dr .da .cc .azd .mt
.dat .vrt .mt.vrt), the cullst inputs read from them, and the cullst
products.  Host code; no GPU is needed."""
import ctypes as C

import numpy as np

from . import track

ST_LSB, ST_MSB = 0, 1
ST_ERR_IO = -6


class OffsetsMeta(C.Structure):
    _fields_ = [("rStart", C.c_int32), ("aStart", C.c_int32), ("deltaR", C.c_int32), ("deltaA", C.c_int32),
                ("scaleFactor", C.c_int32), ("wR", C.c_int32), ("wA", C.c_int32), ("wRa", C.c_int32),
                ("wAa", C.c_int32), ("intGeodat", C.c_char_p), ("imageFile1", C.c_char_p),
                ("imageFile2", C.c_char_p), ("maskFile", C.c_char_p)]


class CullMeta(C.Structure):
    _fields_ = [("rStart", C.c_int32), ("aStart", C.c_int32), ("deltaR", C.c_int32), ("deltaA", C.c_int32),
                ("sR", C.c_int32), ("sA", C.c_int32), ("boxSize", C.c_int32), ("nGood", C.c_int32),
                ("singleMT", C.c_int32), ("maxR", C.c_float), ("maxA", C.c_float), ("corrThresh", C.c_float),
                ("geo1", C.c_char_p), ("geo2", C.c_char_p)]


class CullLSMeta(C.Structure):
    _fields_ = [("fileEarly", C.c_char_p), ("fileLate", C.c_char_p), ("x0", C.c_double), ("y0", C.c_double),
                ("dx", C.c_double), ("dy", C.c_double), ("stepX", C.c_uint32), ("stepY", C.c_uint32),
                ("slowFlag", C.c_int32), ("proj", C.c_int32), ("jdEarly", C.c_double), ("jdLate", C.c_double),
                ("noData", C.c_float)]


_PAR_PATH = 1024
_PAR_STRINGS = ("image1", "image2", "par1", "par2", "vrt1", "vrt2", "initShift", "initialOffsetFile", "outputFile",
                "maskFile", "maskGeodat", "maskVrt", "intFile", "intGeodat", "baseParams")
_PAR_INTS = ("polyShift", "maskFlag", "maskType", "intFlag", "baseFlag", "rStart", "aStart", "deltaR", "deltaA", "nR",
             "nA", "wR", "wA", "wRa", "wAa", "edgePad", "edgePadR", "edgePadA", "scaleFactor", "navgR", "navgA")


class TrackParFile(C.Structure):
    """st_track_par_file: what parseTrack (Strack/parseTrack.c) takes from
    the strack parameter file."""
    _fields_ = [(n, C.c_char * _PAR_PATH) for n in _PAR_STRINGS] + [(n, C.c_int32) for n in _PAR_INTS]


_bound = False


def _lib():
    global _bound
    lib = track.load_library()
    if not _bound:
        vp = C.c_void_p
        lib.st_io_last_error.restype = C.c_char_p
        lib.st_write_offsets.argtypes = [C.c_char_p, C.POINTER(OffsetsMeta), vp, vp, vp, vp, vp, C.c_int32,
                                         C.c_int32, C.c_int32]
        lib.st_read_dat.argtypes = [C.c_char_p, C.POINTER(C.c_int32), C.c_char_p, C.c_char_p, C.c_size_t]
        lib.st_read_offsets.argtypes = [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, vp, vp, vp, vp]
        lib.st_write_cull.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(CullMeta), vp, vp, vp, vp, vp, vp,
                                      C.c_int32, C.c_int32, C.c_int32]
        lib.st_write_cull_ls.argtypes = [C.c_char_p, C.POINTER(CullLSMeta), vp, vp, vp, vp, vp, C.c_int32,
                                         C.c_int32, C.c_int32, C.c_int32, C.c_int32]
        _bound = True
    return lib


class IoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (status %d)" % (msg, code))
        self.code = code


def _check(st):
    if st != 0:
        raise IoError(st, _lib().st_io_last_error().decode())


def _b(s):
    return None if s is None else s.encode()


def _f32(a, shape):
    a = np.ascontiguousarray(a, np.float32)
    if a.shape != shape:
        raise ValueError("grid shape %s, expected %s" % (a.shape, shape))
    return a


def offsets_meta(par, intGeodat, imageFile2, imageFile1=None, maskFile=None):
    """Header fields of an offsets product from the tracking parameters."""
    return OffsetsMeta(par.rStart, par.aStart, par.deltaR, par.deltaA, par.scaleFactor, par.wR, par.wA, par.wRa,
                       par.wAa, _b(intGeodat), _b(imageFile1), _b(imageFile2), _b(maskFile))


def write_offsets(stem, meta, offR, offA, corr, type_, azd=None, byte_order=ST_MSB):
    """stem.dr .da .cc [.azd] .mt .dat .vrt .mt.vrt (sTrackOut.c / writeVrt.c)."""
    nA, nR = np.shape(offR)
    offR, offA, corr = (_f32(a, (nA, nR)) for a in (offR, offA, corr))
    t = np.ascontiguousarray(type_, np.uint8)
    z = None if azd is None else _f32(azd, (nA, nR))
    _check(_lib().st_write_offsets(stem.encode(), C.byref(meta), offR.ctypes.data, offA.ctypes.data,
                                   corr.ctypes.data, None if z is None else z.ctypes.data, t.ctypes.data, nR, nA,
                                   byte_order))


def write_result(stem, meta, res, byte_order=ST_MSB, azd=True):
    """A TrackResult of Device.speckle_track / corr_track_fast."""
    write_offsets(stem, meta, res.offR, res.offA, res.corr, res.type, res.azd if azd else None, byte_order)


def read_dat(path):
    """(r0, a0, nr, na, deltaR, deltaA), geo1, geo2 of a .dat header."""
    hdr = (C.c_int32 * 6)()
    g1 = C.create_string_buffer(1024)
    g2 = C.create_string_buffer(1024)
    _check(_lib().st_read_dat(path.encode(), hdr, g1, g2, 1024))
    return tuple(hdr), g1.value.decode(), g2.value.decode()


def read_offsets(stem, nR, nA, byte_order=ST_MSB):
    """offR, offA, corr, type grids of stem.dr .da .cc .mt (cullst's loader)."""
    offR = np.empty((nA, nR), np.float32)
    offA = np.empty_like(offR)
    corr = np.empty_like(offR)
    t = np.empty((nA, nR), np.uint8)
    _check(_lib().st_read_offsets(stem.encode(), nR, nA, byte_order, offR.ctypes.data, offA.ctypes.data,
                                  corr.ctypes.data, t.ctypes.data))
    return offR, offA, corr, t


def write_cull(out_base, in_base, meta, offRS, offAS, sigmaR, sigmaA, corr, type_, byte_order=ST_MSB):
    """out_base.dr .da .cc .mt .vrt .mt.vrt and in_base.sr .sa (writeCullData.c)."""
    nA, nR = np.shape(offRS)
    arrs = [_f32(a, (nA, nR)) for a in (offRS, offAS, sigmaR, sigmaA, corr)]
    t = np.ascontiguousarray(type_, np.uint8)
    _check(_lib().st_write_cull(out_base.encode(), in_base.encode(), C.byref(meta), *[a.ctypes.data for a in arrs],
                                t.ctypes.data, nR, nA, byte_order))


def write_cull_ls(match_file, meta, XS, YS, sigmaX, sigmaY, type_, n_match, n_attempt, byte_order=ST_MSB):
    """match_file.cull.dx .dy .sx .sy .mtype .dat (Cullls/cullls.c:63-175)."""
    ny, nx = np.shape(XS)
    arrs = [_f32(a, (ny, nx)) for a in (XS, YS, sigmaX, sigmaY)]
    t = np.ascontiguousarray(type_, np.uint8)
    _check(_lib().st_write_cull_ls(match_file.encode(), C.byref(meta), *[a.ctypes.data for a in arrs], t.ctypes.data,
                                   nx, ny, n_match, n_attempt, byte_order))


# ---- inputs before tracking -------------------------------------------------

def _bind_inputs(lib):
    if getattr(lib, "_st_inputs_bound", False):
        return
    lib.st_read_offset_polys.argtypes = [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    i32p, f32p = C.POINTER(C.c_int32), C.POINTER(C.c_float)
    lib.st_read_initial_offsets_header.argtypes = [C.c_char_p, i32p, i32p, i32p, i32p, f32p, f32p]
    lib.st_read_initial_offsets.argtypes = [C.c_char_p, C.c_char_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p]
    lib.st_read_mask_header.argtypes = [C.c_char_p, i32p]
    lib.st_read_mask.argtypes = [C.c_char_p, C.c_int32, C.c_int32, C.c_void_p]
    lib.st_read_interferogram.argtypes = [C.c_char_p, C.c_int32, C.c_int32, C.c_void_p]
    lib.st_read_track_par.argtypes = [C.c_char_p, C.c_int32, C.POINTER(TrackParFile)]
    lib._st_inputs_bound = True


def read_track_par(path, no_complex=False):
    """The strack parameter file (parseTrack, Strack/parseTrack.c:13-190) as
    a dict: file names as str ("" when absent), flags and grid / window
    values as int (rStart, aStart, deltaR, deltaA already divided by
    scaleFactor).  Raises IoError with parseTrack's message when a required
    keyword is missing or a value does not parse."""
    lib = _lib()
    _bind_inputs(lib)
    o = TrackParFile()
    _check(lib.st_read_track_par(path.encode(), 1 if no_complex else 0, C.byref(o)))
    d = {n: getattr(o, n).decode() for n in _PAR_STRINGS}
    d.update({n: getattr(o, n) for n in _PAR_INTS})
    return d


def track_params_from_par(d, **kw):
    """st_track_params from read_track_par's dict (grid, windows, edge pads,
    scale factor, polyShift); the rest from track.default_params / kw."""
    keys = ("rStart", "aStart", "deltaR", "deltaA", "nR", "nA", "wR", "wA", "wRa", "wAa", "edgePadR", "edgePadA",
            "scaleFactor", "polyShift")
    base = {k: d[k] for k in keys}
    base.update(kw)
    return track.default_params(**base)


def read_offset_polys(path):
    """(rShiftPoly, aShiftPoly) of an initial-offset polynomial file."""
    lib = _lib()
    _bind_inputs(lib)
    r = (C.c_double * 3)()
    a = (C.c_double * 3)()
    _check(lib.st_read_offset_polys(path.encode(), r, a))
    return list(r), list(a)


def read_initial_offsets(da_file, dr_file):
    """Initial-offset grids of readBothOffsetsStrack: returns (r0, a0,
    deltaR, deltaA, da, dr) with da / dr as (na, nr) float32 arrays."""
    lib = _lib()
    _bind_inputs(lib)
    r0, a0, nr, na = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
    dR, dA = C.c_float(), C.c_float()
    _check(lib.st_read_initial_offsets_header(da_file.encode(), C.byref(r0), C.byref(a0), C.byref(nr), C.byref(na),
                                              C.byref(dR), C.byref(dA)))
    da = np.empty((na.value, nr.value), np.float32)
    dr = np.empty_like(da)
    _check(lib.st_read_initial_offsets(da_file.encode(), dr_file.encode(), nr.value, na.value, da.ctypes.data,
                                       dr.ctypes.data))
    return r0.value, a0.value, dR.value, dA.value, da, dr


def apply_initial_offsets(par, grids):
    """Point st_track_params at initial-offset grids (polyShift = 0).  The
    arrays must outlive the tracking call; they are returned."""
    from . import abi
    r0, a0, dR, dA, da, dr = grids
    par.polyShift = 0
    par.initDr = abi.ptr(dr, C.c_float)
    par.initDa = abi.ptr(da, C.c_float)
    par.initNa, par.initNr = da.shape
    par.initR0, par.initA0, par.initDeltaR, par.initDeltaA = float(r0), float(a0), float(dR), float(dA)
    return da, dr


def read_mask(dat_file, mask_file):
    """Offset mask (getMask's offset branch): ((r0, a0, nr, na, nrl, nal), mask (na, nr) uint8)."""
    lib = _lib()
    _bind_inputs(lib)
    hdr = (C.c_int32 * 6)()
    _check(lib.st_read_mask_header(dat_file.encode(), hdr))
    r0, a0, nr, na, nrl, nal = tuple(hdr)
    m = np.empty((na, nr), np.uint8)
    _check(lib.st_read_mask(mask_file.encode(), nr, na, m.ctypes.data))
    return (r0, a0, nr, na, nrl, nal), m


def apply_mask(par, hdr, mask):
    """Point st_track_params at a mask read by read_mask; returns the array."""
    from . import abi
    r0, a0, nr, na, nrl, nal = hdr
    par.mask = abi.ptr(mask, C.c_uint8)
    par.maskNr, par.maskNa, par.maskR0, par.maskA0, par.maskNrl, par.maskNal = nr, na, r0, a0, nrl, nal
    return mask


def read_interferogram(path, nr, na):
    """getInt: the (na, 2 nr) float32 (re, im) interferogram of a big-endian
    int16 file; lines past the end of a short file stay 1 + 1i."""
    lib = _lib()
    _bind_inputs(lib)
    intf = np.empty((na, 2 * nr), np.float32)
    _check(lib.st_read_interferogram(path.encode(), nr, na, intf.ctypes.data))
    return intf


def apply_interferogram(par, intf, nrl, nal):
    """Point st_track_params at an interferogram ((na, 2 nr) float32 pairs)
    with nrl x nal looks; returns the array, which must outlive the call."""
    from . import abi
    intf = np.ascontiguousarray(intf, dtype=np.float32)
    par.intf = abi.ptr(intf, C.c_float)
    par.intNa, par.intNr = intf.shape[0], intf.shape[1] // 2
    par.intNrl, par.intNal = nrl, nal
    return intf


# ---- GDAL VRT inputs (raw bands) ---------------------------------------------

def _bind_vrt(lib):
    if getattr(lib, "_st_vrt_bound", False):
        return
    vp = C.c_void_p
    lib.st_vrt_last_error.restype = C.c_char_p
    lib.st_vrt_open.argtypes = [C.c_char_p, C.POINTER(vp)]
    lib.st_vrt_close.argtypes = [vp]
    lib.st_vrt_close.restype = None
    i32p = C.POINTER(C.c_int32)
    lib.st_vrt_size.argtypes = [vp, i32p, i32p, i32p]
    lib.st_vrt_metadata.argtypes = [vp, C.c_char_p, C.c_char_p, C.c_size_t]
    lib.st_vrt_band_description.argtypes = [vp, C.c_int32, C.c_char_p, C.c_size_t]
    for nm in ("st_vrt_read_f32", "st_vrt_read_u8", "st_vrt_read_c64"):
        getattr(lib, nm).argtypes = [vp, C.c_int32, vp]
    lib._st_vrt_bound = True


class Vrt:
    """A VRT with raw bands: .nx, .ny, .n_bands, .metadata(key),
    .description(band), .read(band, 'f32' | 'u8' | 'c64')."""

    def __init__(self, path):
        self._lib = _lib()
        _bind_vrt(self._lib)
        self._h = C.c_void_p()
        self._ck(self._lib.st_vrt_open(path.encode(), C.byref(self._h)))
        nx, ny, nb = C.c_int32(), C.c_int32(), C.c_int32()
        self._ck(self._lib.st_vrt_size(self._h, C.byref(nx), C.byref(ny), C.byref(nb)))
        self.nx, self.ny, self.n_bands = nx.value, ny.value, nb.value

    def _ck(self, st):
        if st != 0:
            raise IoError(st, self._lib.st_vrt_last_error().decode())

    def metadata(self, key):
        buf = C.create_string_buffer(4096)
        self._ck(self._lib.st_vrt_metadata(self._h, key.encode(), buf, len(buf)))
        return buf.value.decode()

    def description(self, band):
        buf = C.create_string_buffer(4096)
        self._ck(self._lib.st_vrt_band_description(self._h, band, buf, len(buf)))
        return buf.value.decode()

    def read(self, band, kind="f32"):
        if kind == "f32":
            out = np.empty((self.ny, self.nx), np.float32)
        elif kind == "u8":
            out = np.empty((self.ny, self.nx), np.uint8)
        elif kind == "c64":
            out = np.empty((self.ny, self.nx), np.complex64)
        else:
            raise ValueError(kind)
        self._ck(getattr(self._lib, "st_vrt_read_" + kind)(self._h, band, out.ctypes.data))
        return out

    def close(self):
        if self._h:
            self._lib.st_vrt_close(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def _atoi(s):
    """C atoi: the leading integer of s, 0 when there is none."""
    import re
    m = re.match(r"\s*([+-]?\d+)", s)
    return int(m.group(1)) if m else 0


def _atof(s):
    """C atof: the leading floating-point number of s, 0 when there is none."""
    import re
    m = re.match(r"\s*([+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?)", s)
    return float(m.group(1)) if m else 0.0


def read_initial_offsets_vrt(vrt_path):
    """readBothOffsetsStrackVrt (Strack/readBothOffsetsStrack.c:29-76): the
    (r0, a0, deltaR, deltaA, da, dr) tuple of read_initial_offsets from an
    offsets VRT (band 1 range, band 2 azimuth offsets, checked by their
    descriptions)."""
    with Vrt(vrt_path) as v:
        r0, a0 = _atoi(v.metadata("r0")), _atoi(v.metadata("a0"))
        dR, dA = _atof(v.metadata("deltaR")), _atof(v.metadata("deltaA"))
        if "Range" not in v.description(1):
            raise IoError(-6, "Problem reading range offsets from %s" % vrt_path)
        dr = v.read(1)
        if "Azimuth" not in v.description(2):
            raise IoError(-6, "Problem reading azimuth offsets from %s" % vrt_path)
        da = v.read(2)
    return r0, a0, dR, dA, da, dr


def read_mask_vrt(vrt_path):
    """getMaskVRT (Strack/getMask.c:22-48): ((r0, a0, nr, na, nrl, nal), mask)
    with nrl / nal the integer deltaR / deltaA of the metadata."""
    with Vrt(vrt_path) as v:
        hdr = (_atoi(v.metadata("r0")), _atoi(v.metadata("a0")), v.nx, v.ny, _atoi(v.metadata("deltaR")),
               _atoi(v.metadata("deltaA")))
        return hdr, v.read(1, "u8")


def read_cull_vrt(vrt_path, mt_vrt_path):
    """loadFromVRT (Cullst/loadCullData.c:46-92): metadata r0 a0 deltaR deltaA
    and the offR, offA, corr bands of the offsets VRT, the match types of
    the .mt VRT."""
    with Vrt(vrt_path) as v:
        meta = dict(r0=_atoi(v.metadata("r0")), a0=_atoi(v.metadata("a0")), deltaR=_atof(v.metadata("deltaR")),
                    deltaA=_atof(v.metadata("deltaA")))
        offR, offA, corr = v.read(1), v.read(2), v.read(3)
    with Vrt(mt_vrt_path) as m:
        typ = m.read(1, "u8")
    return meta, offR, offA, corr, typ
