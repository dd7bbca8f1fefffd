"""The files either side of the tracking path (include/speckle_io.h): what
strack / strackw write after tracking, what cullst reads, what cullst
writes.  Host code, no GPU.

Pinned by: numpy's big-endian / little-endian float32 encodings for the raw
rasters (the reference's fwriteBS writes each float32 big-endian on a
little-endian host, Strack/sTrackOut.c:11-15); the reference's printf
formats for the .dat header (Strack/speckleTrack.c:1453-1468); the file
names of Cullst/cullst.c:240-300.  The VRT XML is checked for the GDAL
VRTRawRasterBand fields GDAL needs; its exact bytes come from GDAL's writer
in the reference (writeSingleVRT, not vendored) and are parity unpinned."""
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from specklesource_amd import files

NR, NA = 37, 23


def _grids(seed=0):
    rng = np.random.default_rng(seed)
    offR = rng.normal(0, 5, (NA, NR)).astype(np.float32)
    offA = rng.normal(0, 5, (NA, NR)).astype(np.float32)
    offA[3, 4] = np.float32(-2e9)
    offR[3, 4] = np.float32(-2e9)
    corr = rng.uniform(0, 1, (NA, NR)).astype(np.float32)
    azd = rng.normal(0, 1, (NA, NR)).astype(np.float32)
    t = rng.integers(0, 4, (NA, NR)).astype(np.uint8)
    return offR, offA, corr, azd, t


def _meta():
    return files.OffsetsMeta(rStart=100, aStart=200, deltaR=24, deltaA=12, scaleFactor=2, wR=64, wA=64, wRa=48,
                             wAa=48, intGeodat=b"geodat10x2.in", imageFile1=b"/data/p1/slc.1",
                             imageFile2=b"/data/p2/slc.2", maskFile=None)


@pytest.mark.parametrize("order,dt", [(files.ST_MSB, ">f4"), (files.ST_LSB, "<f4")])
def test_offsets_rasters_byte_exact(tmp_path, order, dt):
    offR, offA, corr, azd, t = _grids()
    stem = str(tmp_path / "offsets")
    files.write_offsets(stem, _meta(), offR, offA, corr, t, azd=azd, byte_order=order)
    for ext, a in ((".dr", offR), (".da", offA), (".cc", corr), (".azd", azd)):
        raw = open(stem + ext, "rb").read()
        assert raw == a.astype(dt).tobytes(), ext
    assert open(stem + ".mt", "rb").read() == t.tobytes()


def test_strackw_writes_no_azd(tmp_path):
    offR, offA, corr, _, t = _grids(1)
    stem = str(tmp_path / "w")
    files.write_offsets(stem, _meta(), offR, offA, corr, t, azd=None)
    assert not os.path.exists(stem + ".azd")
    assert os.path.getsize(stem + ".dr") == 4 * NR * NA


def test_dat_header_text(tmp_path):
    offR, offA, corr, _, t = _grids()
    stem = str(tmp_path / "offsets")
    files.write_offsets(stem, _meta(), offR, offA, corr, t)
    # "%i %i %i %i %i %i\n" of r0*sf a0*sf nR nA dR*sf dA*sf, then
    # "geodat dirname(image2)/geodat"
    want = "%i %i %i %i %i %i\n" % (200, 400, NR, NA, 48, 24) + "geodat10x2.in /data/p2/geodat10x2.in\n"
    assert open(stem + ".dat").read() == want
    hdr, g1, g2 = files.read_dat(stem + ".dat")
    assert hdr == (200, 400, NR, NA, 48, 24)
    assert (g1, g2) == ("geodat10x2.in", "/data/p2/geodat10x2.in")


def test_read_dat_without_geodat_line(tmp_path):
    p = tmp_path / "x.dat"
    p.write_text("; comment\n1 2 3 4 5 6\n")
    hdr, g1, g2 = files.read_dat(str(p))
    assert hdr == (1, 2, 3, 4, 5, 6) and g1 == "" and g2 == ""
    p.write_text("1 2 3\n")
    with pytest.raises(files.IoError) as e:
        files.read_dat(str(p))
    assert e.value.code == files.ST_ERR_IO


def _vrt(path):
    root = ET.parse(path).getroot()
    meta = [(m.get("key"), m.text) for m in root.find("Metadata")]
    bands = root.findall("VRTRasterBand")
    return root, meta, bands


@pytest.mark.parametrize("order", [files.ST_MSB, files.ST_LSB])
def test_offsets_vrt(tmp_path, order):
    offR, offA, corr, _, t = _grids()
    stem = str(tmp_path / "offsets")
    files.write_offsets(stem, _meta(), offR, offA, corr, t, byte_order=order)
    root, meta, bands = _vrt(stem + ".vrt")
    assert (root.get("rasterXSize"), root.get("rasterYSize")) == (str(NR), str(NA))
    keys = [k for k, _ in meta]
    assert keys == ["r0", "a0", "deltaA", "deltaR", "sigmaStreaks", "sigmaRange", "geo1", "geo2", "Image1",
                    "Image2", "mask", "wR", "wA", "wRa", "wAa", "scaleFactor", "ByteOrder"]
    md = dict(meta)
    assert md["r0"] == "200" and md["deltaA"] == "24" and md["sigmaStreaks"] == "0.000000"
    assert md["geo2"] == "/data/p2/geodat10x2.in" and md["mask"] == "(null)"
    assert md["ByteOrder"] == ("MSB" if order == files.ST_MSB else "LSB")
    assert [b.find("Description").text for b in bands] == ["RangeOffsets", "AzimuthOffsets", "Correlation"]
    for b, ext in zip(bands, (".dr", ".da", ".cc")):
        assert b.get("dataType") == "Float32" and b.get("subClass") == "VRTRawRasterBand"
        assert b.find("SourceFilename").text == "offsets" + ext
        assert b.find("SourceFilename").get("relativeToVRT") == "1"
        assert b.find("NoDataValue").text == "-2000000000"
        assert b.find("LineOffset").text == str(4 * NR) and b.find("PixelOffset").text == "4"
        assert b.find("ByteOrder").text == md["ByteOrder"]
    root, meta, bands = _vrt(stem + ".mt.vrt")
    assert "ByteOrder" not in dict(meta)
    assert len(bands) == 1 and bands[0].get("dataType") == "Byte" and bands[0].find("NoDataValue") is None
    assert bands[0].find("SourceFilename").text == "offsets.mt"


def test_read_offsets_round_trip_and_type_rule(tmp_path):
    offR, offA, corr, _, t = _grids(2)
    offA[5, 6] = np.float32(-2.5e9)  # below (float)(-LARGEINT + 1): type forced to 0
    t[5, 6] = 2
    t[3, 4] = 2  # the -2e9 sentinel itself is not below the bound
    stem = str(tmp_path / "o")
    files.write_offsets(stem, _meta(), offR, offA, corr, t)
    r, a, c, tt = files.read_offsets(stem, NR, NA)
    assert r.tobytes() == offR.tobytes() and a.tobytes() == offA.tobytes() and c.tobytes() == corr.tobytes()
    want = t.copy()
    want[5, 6] = 0
    assert np.array_equal(tt, want)
    with pytest.raises(files.IoError):
        files.read_offsets(stem, NR + 1, NA)  # short read


def test_cull_products(tmp_path):
    offR, offA, corr, azd, t = _grids(3)
    sr = np.abs(azd)
    sa = np.abs(azd) * 2
    out = str(tmp_path / "out" / "culled")
    os.makedirs(os.path.dirname(out))
    inb = str(tmp_path / "offsets")
    m = files.CullMeta(rStart=200, aStart=400, deltaR=48, deltaA=24, sR=3, sA=3, boxSize=9, nGood=20, singleMT=0,
                       maxR=1.5, maxA=2.0, corrThresh=0.07, geo1=b"g1.in", geo2=b"/x/g1.in")
    files.write_cull(out, inb, m, offR, offA, sr, sa, corr, t)
    for path, a in ((out + ".dr", offR), (out + ".da", offA), (inb + ".sr", sr), (inb + ".sa", sa),
                    (out + ".cc", corr)):
        assert open(path, "rb").read() == a.astype(">f4").tobytes(), path
    assert open(out + ".mt", "rb").read() == t.tobytes()
    root, meta, bands = _vrt(out + ".vrt")
    keys = [k for k, _ in meta]
    assert keys == ["r0", "a0", "deltaA", "deltaR", "sigmaStreaks", "sigmaRange", "sR", "sA", "boxSize", "nGood",
                    "maxR", "maxA", "corrThresh", "singleMT", "geo1", "geo2", "ByteOrder"]
    assert dict(meta)["corrThresh"] == "0.070000" and dict(meta)["maxR"] == "1.500000"
    assert [b.find("Description").text for b in bands] == ["RangeOffsets", "AzimuthOffsets", "RangeSigma",
                                                           "AzimuthSigma", "Correlation"]
    src = [(b.find("SourceFilename").text, b.find("SourceFilename").get("relativeToVRT")) for b in bands]
    assert src[0] == ("culled.dr", "1") and src[2] == (inb + ".sr", "0")


def test_bad_arguments():
    offR, offA, corr, _, t = _grids()
    with pytest.raises(files.IoError) as e:
        files.write_offsets("/nonexistent-dir/x", _meta(), offR, offA, corr, t)
    assert e.value.code == files.ST_ERR_IO
    with pytest.raises(files.IoError) as e:
        files.write_offsets("/tmp/x", _meta(), offR, offA, corr, t, byte_order=7)
    assert e.value.code == -1


# ---- inputs before tracking -------------------------------------------------
# Formats from the reference's own scanf/fread calls: parseInitialOffsets.c
# (keyword lines, "%lf%lf%lf"), readBothOffsetsStrack.c:94-120 (file.dat
# "%i%i%i%i%f%f", big-endian float32 grids), getMask.c:75-118 (the same
# header, nrl / nal = (int) deltaR / deltaA, bytes).

def test_offset_polys(tmp_path):
    p = tmp_path / "init.off"
    p.write_text("; initial offsets\nsome_other_key: 1 2 3\nazimuth_offset_polynomial: -1.25 2e-4 -3.5e-6\n"
                 "range_offset_polynomial:  4.5 -1e-3 7e-7\n")
    r, a = files.read_offset_polys(str(p))
    assert r == [4.5, -1e-3, 7e-7] and a == [-1.25, 2e-4, -3.5e-6]
    p.write_text("range_offset_polynomial: 1 2 3\n")
    with pytest.raises(files.IoError):
        files.read_offset_polys(str(p))
    p.write_text("range_offset_polynomial: 1 2\nazimuth_offset_polynomial: 1 2 3\n")
    with pytest.raises(files.IoError):
        files.read_offset_polys(str(p))


def test_initial_offset_grids(tmp_path):
    rng = np.random.default_rng(5)
    da = rng.normal(0, 3, (9, 13)).astype(np.float32)
    dr = rng.normal(0, 3, (9, 13)).astype(np.float32)
    base = tmp_path / "prior.da"
    (tmp_path / "prior.da.dat").write_text("120 240 13 9 40.0 80.0\n")
    base.write_bytes(da.astype(">f4").tobytes())
    (tmp_path / "prior.dr").write_bytes(dr.astype(">f4").tobytes())
    r0, a0, dR, dA, gda, gdr = files.read_initial_offsets(str(base), str(tmp_path / "prior.dr"))
    assert (r0, a0, dR, dA) == (120, 240, 40.0, 80.0)
    assert gda.tobytes() == da.tobytes() and gdr.tobytes() == dr.tobytes()
    from specklesource_amd import track
    p = track.default_params(nR=4, nA=4, wR=32, wA=32, wRa=32, wAa=32)
    keep = files.apply_initial_offsets(p, (r0, a0, dR, dA, gda, gdr))
    assert p.polyShift == 0 and (p.initNr, p.initNa) == (13, 9) and p.initDeltaA == 80.0 and keep[0] is gda


def test_offset_mask(tmp_path):
    rng = np.random.default_rng(6)
    m = rng.integers(0, 2, (7, 11)).astype(np.uint8)
    (tmp_path / "mask.dat").write_text("; mask\n10 20 11 7 12.0 3.0\n")
    (tmp_path / "mask.bytes").write_bytes(m.tobytes())
    hdr, gm = files.read_mask(str(tmp_path / "mask.dat"), str(tmp_path / "mask.bytes"))
    assert hdr == (10, 20, 11, 7, 12, 3) and np.array_equal(gm, m)
    (tmp_path / "short.bytes").write_bytes(m.tobytes()[:-1])
    with pytest.raises(files.IoError):
        files.read_mask(str(tmp_path / "mask.dat"), str(tmp_path / "short.bytes"))


def test_interferogram(tmp_path):
    """getInt (Strack/getInt.c:46-75): big-endian int16 (re, im) pairs; a
    short file leaves the missing lines at 1 + 1i."""
    rng = np.random.default_rng(7)
    nr, na = 9, 6
    z = rng.integers(-30000, 30000, (na, 2 * nr)).astype(">i2")
    (tmp_path / "full.int").write_bytes(z.tobytes())
    got = files.read_interferogram(str(tmp_path / "full.int"), nr, na)
    assert got.dtype == np.float32 and np.array_equal(got, z.astype(np.float32))
    (tmp_path / "short.int").write_bytes(z[:4].tobytes() + z[4, :5].tobytes())
    got = files.read_interferogram(str(tmp_path / "short.int"), nr, na)
    assert np.array_equal(got[:4], z[:4].astype(np.float32)) and np.all(got[4:] == 1.0)
    from specklesource_amd import track
    p = track.default_params(nR=4, nA=4, wR=32, wA=32, wRa=32, wAa=32)
    keep = files.apply_interferogram(p, got, 2, 8)
    assert (p.intNr, p.intNa, p.intNrl, p.intNal) == (nr, na, 2, 8) and keep is got
    with pytest.raises(files.IoError):
        files.read_interferogram(str(tmp_path / "missing.int"), nr, na)


def test_cull_ls_products(tmp_path):
    """writeLSCulledOffsets (Cullls/cullls.c:63-175): four big-endian float32
    rasters, the type bytes, and the .cull.dat header with the reference's
    formats (its stepY line carries stepX and the other way round)."""
    rng = np.random.default_rng(8)
    ny, nx = 7, 9
    XS = rng.normal(0, 2, (ny, nx)).astype(np.float32)
    YS = rng.normal(0, 2, (ny, nx)).astype(np.float32)
    sx = rng.uniform(0.1, 0.5, (ny, nx)).astype(np.float32)
    sy = rng.uniform(0.1, 0.5, (ny, nx)).astype(np.float32)
    XS[2, 3] = YS[2, 3] = np.float32(-2e9)
    t = rng.integers(0, 255, (ny, nx)).astype(np.uint8)
    m = files.CullLSMeta(fileEarly=b"LC08_early.TIF", fileLate=b"LC08_late.TIF", x0=-3001234.5, y0=123456.25,
                         dx=15.0, dy=15.0, stepX=24, stepY=12, slowFlag=1, proj=3413, jdEarly=2459000.5,
                         jdLate=2459016.5, noData=-2e9)
    base = str(tmp_path / "pair.match")
    files.write_cull_ls(base, m, XS, YS, sx, sy, t, 60, 63)
    for ext, a in ((".cull.dx", XS), (".cull.dy", YS), (".cull.sx", sx), (".cull.sy", sy)):
        assert (tmp_path / ("pair.match" + ext)).read_bytes() == a.astype(">f4").tobytes()
    assert (tmp_path / "pair.match.cull.mtype").read_bytes() == t.tobytes()
    good = np.ones((ny, nx), bool)
    good[2, 3] = False
    rx = np.sqrt(np.sum((sx[good] * sx[good]).astype(np.float64)) / good.sum())
    ry = np.sqrt(np.sum((sy[good] * sy[good]).astype(np.float64)) / good.sum())
    want = ("fileEarly = LC08_early.TIF\nfileLate = LC08_late.TIF\nx0 = %11.2f\ny0 = %11.2f\ndx = %10.5f\n"
            "dy = %10.5f\nstepY = 24\nstepX = 12\nnx = %9d\nny = %9d\nslowFlag = %9d\nEPSG = %9d\n"
            "earlyImageJD = %10f\nlateImageJD = %10f\nIntervalBetweenImages = %5d\n"
            "Success_rate_for_attempted_matches(%%) =  %7.2f \nCulled_rate_for_attempted_matches(%%) =  %7.2f \n"
            "Mean_sigmaX = %11.3f\nMean_sigmaY = %11.3f\n& \n") % (
        -3001234.5, 123456.25, 15.0, 15.0, nx, ny, 1, 3413, 2459000.5, 2459016.5, 16,
        float(np.float32(np.float32(60) / np.float32(63))) * 100.0,
        float(np.float32(np.float32(good.sum()) / np.float32(63))) * 100.0, rx, ry)
    assert (tmp_path / "pair.match.cull.dat").read_text() == want
