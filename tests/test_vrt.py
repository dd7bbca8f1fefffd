"""GDAL VRT inputs with raw bands (include/speckle_io.h st_vrt_*): the readers
behind readBothOffsetsStrackVrt, getMaskVRT and loadFromVRT.  Host code, no
GPU.

Pinned by: round trips through this library's own VRT writers; hand-written
VRTs in the VRTRawRasterBand form of GDAL's VRT format description
(band-interleaved-by-pixel offsets, an image offset, both byte orders, the
defaults for omitted offsets and byte order, an absolute source path, an XML
declaration and comments), whose expected arrays come from numpy's own
encodings of the same bytes.  How GDAL itself reads other band kinds is not
covered (they are refused)."""

import numpy as np
import pytest

from specklesource_amd import files

NR, NA = 13, 9


def _vrt(path, body, nx=NR, ny=NA, meta=None, decl=True):
    md = "".join('    <MDI key="%s">%s</MDI>\n' % kv for kv in (meta or {}).items())
    x = ('<?xml version="1.0"?>\n' if decl else "") + '<VRTDataset rasterXSize="%d" rasterYSize="%d">\n' % (nx, ny)
    x += "  <!-- written by hand -->\n  <Metadata>\n" + md + "  </Metadata>\n" + body + "</VRTDataset>\n"
    path.write_text(x)
    return str(path)


def _band(n, dtype, src, rel=True, image=None, pixel=None, line=None, order=None, desc=None):
    s = '  <VRTRasterBand dataType="%s" band="%d" subClass="VRTRawRasterBand">\n' % (dtype, n)
    if desc:
        s += "    <Description>%s</Description>\n" % desc
    s += '    <SourceFilename relativeToVRT="%d">%s</SourceFilename>\n' % (1 if rel else 0, src)
    for tag, v in (("ImageOffset", image), ("PixelOffset", pixel), ("LineOffset", line), ("ByteOrder", order)):
        if v is not None:
            s += "    <%s>%s</%s>\n" % (tag, v, tag)
    return s + "  </VRTRasterBand>\n"


def test_interleaved_bands_offsets_and_orders(tmp_path):
    rng = np.random.default_rng(1)
    a = rng.normal(0, 3, (NA, NR)).astype(np.float32)
    b = rng.normal(0, 3, (NA, NR)).astype(np.float32)
    # band-interleaved-by-pixel, big-endian, after a 16-byte header
    bip = np.empty((NA, NR, 2), ">f4")
    bip[..., 0] = a
    bip[..., 1] = b
    (tmp_path / "bip.raw").write_bytes(b"H" * 16 + bip.tobytes())
    i16 = rng.integers(-30000, 30000, (NA, NR)).astype("<i2")
    (tmp_path / "i16.raw").write_bytes(i16.tobytes())
    u8 = rng.integers(0, 255, (NA, NR)).astype(np.uint8)
    (tmp_path / "u8.raw").write_bytes(u8.tobytes())
    body = (_band(1, "Float32", "bip.raw", image=16, pixel=8, line=8 * NR, order="MSB", desc="first")
            + _band(2, "Float32", "bip.raw", image=20, pixel=8, line=8 * NR, order="MSB", desc="second")
            + _band(3, "Int16", str(tmp_path / "i16.raw"), rel=False)  # defaults: 2, 2 NR, LSB
            + _band(4, "Byte", "u8.raw"))
    p = _vrt(tmp_path / "x.vrt", body, meta={"r0": "12", "note": "a &amp; b"})
    with files.Vrt(p) as v:
        assert (v.nx, v.ny, v.n_bands) == (NR, NA, 4)
        assert v.metadata("r0") == "12" and v.metadata("note") == "a & b"
        assert v.description(1) == "first" and v.description(3) == ""
        assert np.array_equal(v.read(1), a) and np.array_equal(v.read(2), b)
        assert np.array_equal(v.read(3), i16.astype(np.float32))
        assert np.array_equal(v.read(4, "u8"), u8) and np.array_equal(v.read(4), u8.astype(np.float32))
        with pytest.raises(files.IoError):
            v.read(1, "u8")
        with pytest.raises(files.IoError):
            v.metadata("missing")
        with pytest.raises(files.IoError):
            v.read(5)


def test_complex_bands(tmp_path):
    rng = np.random.default_rng(2)
    z = (rng.normal(0, 100, (NA, NR)) + 1j * rng.normal(0, 100, (NA, NR))).astype(np.complex64)
    (tmp_path / "slc.c64").write_bytes(z.astype(">c8").tobytes())
    q = np.empty((NA, NR, 2), ">i2")
    q[..., 0] = np.round(z.real)
    q[..., 1] = np.round(z.imag)
    (tmp_path / "slc.ci16").write_bytes(q.tobytes())
    body = _band(1, "CFloat32", "slc.c64", order="MSB") + _band(2, "CInt16", "slc.ci16", order="MSB")
    with files.Vrt(_vrt(tmp_path / "slc.vrt", body, decl=False)) as v:
        assert np.array_equal(v.read(1, "c64"), z)
        assert np.array_equal(v.read(2, "c64"), (q[..., 0] + 1j * q[..., 1]).astype(np.complex64))
        with pytest.raises(files.IoError):
            v.read(1, "f32")


def test_refused_and_malformed(tmp_path):
    (tmp_path / "r.raw").write_bytes(bytes(4 * NR * NA))
    simple = '  <VRTRasterBand dataType="Float32" band="1">\n    <SimpleSource/>\n  </VRTRasterBand>\n'
    with pytest.raises(files.IoError, match="VRTRawRasterBand"):
        files.Vrt(_vrt(tmp_path / "s.vrt", simple))
    (tmp_path / "bad.vrt").write_text('<VRTDataset rasterXSize="3" rasterYSize="2"><Metadata></VRTDataset>')
    with pytest.raises(files.IoError):
        files.Vrt(str(tmp_path / "bad.vrt"))
    short = _vrt(tmp_path / "short.vrt", _band(1, "Float32", "r.raw"), nx=NR, ny=NA + 1)
    with files.Vrt(short) as v, pytest.raises(files.IoError, match="short read"):
        v.read(1)


@pytest.mark.parametrize("order", [files.ST_MSB, files.ST_LSB])
def test_offsets_vrt_round_trip(tmp_path, order):
    """st_write_offsets' .vrt read back as readBothOffsetsStrackVrt reads it."""
    from tests.test_files import _grids, _meta
    offR, offA, corr, azd, t = _grids(3)
    stem = str(tmp_path / "offsets")
    files.write_offsets(stem, _meta(), offR, offA, corr, t, azd=azd, byte_order=order)
    r0, a0, dR, dA, da, dr = files.read_initial_offsets_vrt(stem + ".vrt")
    # the headers record full-resolution positions: rStart, aStart, deltaR, deltaA x scaleFactor 2
    assert (r0, a0, dR, dA) == (200, 400, 48.0, 24.0)
    assert np.array_equal(dr, offR) and np.array_equal(da, offA)
    with files.Vrt(stem + ".vrt") as v:
        assert [v.description(b) for b in range(1, v.n_bands + 1)][:3] == ["RangeOffsets", "AzimuthOffsets",
                                                                       "Correlation"]
        assert np.array_equal(v.read(3), corr)


def test_offsets_vrt_band_check(tmp_path):
    """Band 1 must be range offsets (readBothOffsetsStrack.c:62-63)."""
    a = np.zeros((NA, NR), np.float32)
    (tmp_path / "a.raw").write_bytes(a.tobytes())
    body = _band(1, "Float32", "a.raw", desc="AzimuthOffsets") + _band(2, "Float32", "a.raw", desc="RangeOffsets")
    p = _vrt(tmp_path / "o.vrt", body, meta={"r0": "1", "a0": "2", "deltaR": "3", "deltaA": "4"})
    with pytest.raises(files.IoError, match="range offsets"):
        files.read_initial_offsets_vrt(p)


def test_mask_vrt(tmp_path):
    rng = np.random.default_rng(4)
    m = rng.integers(0, 2, (NA, NR)).astype(np.uint8)
    (tmp_path / "mask.bytes").write_bytes(m.tobytes())
    p = _vrt(tmp_path / "mask.vrt", _band(1, "Byte", "mask.bytes"),
             meta={"r0": "10", "a0": "20", "deltaR": "12.0", "deltaA": "3"})
    hdr, gm = files.read_mask_vrt(p)
    assert hdr == (10, 20, NR, NA, 12, 3) and np.array_equal(gm, m)


def test_cull_vrt_round_trip(tmp_path):
    """write_cull's products are cullst outputs; loadFromVRT reads the
    tracking products (st_write_offsets) as cullst inputs."""
    from tests.test_files import _grids, _meta
    offR, offA, corr, azd, t = _grids(5)
    stem = str(tmp_path / "track")
    files.write_offsets(stem, _meta(), offR, offA, corr, t, azd=azd)
    meta, gR, gA, gC, gT = files.read_cull_vrt(stem + ".vrt", stem + ".mt.vrt")
    assert meta == dict(r0=200, a0=400, deltaR=48.0, deltaA=24.0)
    assert np.array_equal(gR, offR) and np.array_equal(gA, offA) and np.array_equal(gC, corr)
    assert np.array_equal(gT, t)
