"""The reference-side C bindings (specklesource_amd/cbind: speckleTrack,
corrTrackFast and the four cull stages under the reference's names and
struct fields) and the strack / strackw / cullst programs built on the C
ABI.

CPU: they compile and link against the library with plain gcc, and the
programs check their arguments as the reference's readArgs do.  GPU: the
plain-C test drives the bindings on SLC files (closed-form shifts), and
the programs' products are byte-identical to the library's own results
written by the same writers."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from specklesource_amd import abi, configs, files, track
from tests.helpers import periodic_pair, small_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CBIND = os.path.join(ROOT, "specklesource_amd", "cbind")
BUILD = os.path.join(CBIND, "_build")


def _built():
    subprocess.run(["make", "-s", "-C", CBIND], check=True)
    return {n: os.path.join(BUILD, n) for n in ("test_bindings", "strack", "strackw", "cullst")}


def test_bindings_and_programs_build_and_check_arguments():
    exe = _built()
    for n, p in exe.items():
        assert os.access(p, os.X_OK), n
    # argument errors stop before any GPU call, with the reference's usage
    r = subprocess.run([exe["strack"]], capture_output=True, text=True)
    assert r.returncode == 1 and "-noComplex -singleAmp" in r.stderr
    r = subprocess.run([exe["strack"], "-bogus", "x.par"], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr
    r = subprocess.run([exe["strackw"], "-noComplex", "x.par"], capture_output=True, text=True)
    assert r.returncode == 1 and "-integerComplex -LSB" in r.stderr
    r = subprocess.run([exe["cullst"], "-corrThresh", "1.0", "in", "out"], capture_output=True, text=True)
    assert r.returncode == 1 and "coorThresh must be in interval" in r.stderr
    r = subprocess.run([exe["cullst"], "-useSim", "sim", "in", "out"], capture_output=True, text=True)
    assert r.returncode == 1 and "useSim" in r.stderr


@pytest.mark.gpu
def test_plain_c_test_drives_the_bindings(tmp_path):
    exe = _built()
    a, b = periodic_pair(32, 256, 256, 3, -2, seed=5)
    (tmp_path / "p32_1.slc").write_bytes(a.astype(">c8").tobytes())
    (tmp_path / "p32_2.slc").write_bytes(b.astype(">c8").tobytes())
    a, b = periodic_pair(64, 400, 400, -5, 4, seed=6)
    (tmp_path / "p64_1.slc").write_bytes(a.astype(">c8").tobytes())
    (tmp_path / "p64_2.slc").write_bytes(b.astype(">c8").tobytes())
    r = subprocess.run([exe["test_bindings"], str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "bindings ok" in r.stdout


def _slc_with_vrt(path, img):
    """MSB complex64 raster and the VRT beside it (<image>.vrt)."""
    path.write_bytes(img.astype(">c8").tobytes())
    na, nr = img.shape
    (path.parent / (path.name + ".vrt")).write_text(
        '<VRTDataset rasterXSize="%d" rasterYSize="%d">\n'
        '  <VRTRasterBand dataType="CFloat32" band="1" subClass="VRTRawRasterBand">\n'
        '    <SourceFilename relativeToVRT="1">%s</SourceFilename>\n'
        "    <ByteOrder>MSB</ByteOrder>\n"
        "  </VRTRasterBand>\n</VRTDataset>\n" % (nr, na, path.name))


RASTERS = (".dr", ".da", ".cc", ".mt")


@pytest.mark.gpu
@pytest.mark.parametrize("prog", ["strack", "strackw"])
def test_programs_write_what_the_library_writes(device, tmp_path, prog):
    exe = _built()
    s = configs._spec(480, 480, 71, dr0=1.3, da0=-0.7, gmin=0.6, gmax=0.9)
    a, b = track.synth_host(s)
    _slc_with_vrt(tmp_path / "a.slc", a)
    _slc_with_vrt(tmp_path / "b.slc", b)
    (tmp_path / "zero.base").write_text("; baseline\n0.0 0.0 0.0 0.0\n")
    (tmp_path / "init.off").write_text("range_offset_polynomial: 1 0 0\nazimuth_offset_polynomial: -1 0 0\n")
    w = 32 if prog == "strack" else 64
    p = small_params(w, 480, 480, step=w // 2, rShiftPoly=[1, 0, 0], aShiftPoly=[-1, 0, 0])
    navg = 2 if prog == "strackw" else 1
    extra = "edgepad = 8\nscalefactor = 1\nnavgR = %d\nnavgA = %d\n" % (navg, navg)
    (tmp_path / "track.par").write_text(
        "image1 = {d}/a.slc\nimage2 = {d}/b.slc\ninitialoffsetfile = {d}/init.off\noutputfile = {d}/out\n"
        "baseparams = {d}/zero.base\nintgeodat = {d}/int.geodat\n"
        "rstart = {r}\nastart = {a}\ndeltar = {dr}\ndeltaa = {da}\nnR = {nr}\nnA = {na}\nwr = {w}\nwa = {w}\n"
        "wra = {w}\nwaa = {w}\n{extra}".format(d=tmp_path, r=p.rStart, a=p.aStart, dr=p.deltaR, da=p.deltaA,
                                                nr=p.nR, na=p.nA, w=w, extra=extra))
    run = subprocess.run([exe[prog], str(tmp_path / "track.par")], capture_output=True, text=True, timeout=120)
    print(run.stderr)
    assert run.returncode == 0, run.stderr
    # the library on the same inputs, written by the same writer
    d = files.read_track_par(str(tmp_path / "track.par"))
    q = files.track_params_from_par(d, rShiftPoly=[1, 0, 0], aShiftPoly=[-1, 0, 0], navgR=d["navgR"],
                                    navgA=d["navgA"], noComplex=1 if prog == "strackw" else 0)
    pid = device.upload_pair(a, b)
    try:
        res = device.speckle_track(pid, q) if prog == "strack" else device.corr_track_fast(pid, q)
    finally:
        device.release_pair(pid)
    meta = files.offsets_meta(q, str(tmp_path / "int.geodat"), str(tmp_path / "b.slc"),
                              imageFile1=str(tmp_path / "a.slc"))
    files.write_result(str(tmp_path / "lib"), meta, res, azd=prog == "strack")
    exts = RASTERS + ((".azd",) if prog == "strack" else ())
    for e in exts:
        assert (tmp_path / ("out" + e)).read_bytes() == (tmp_path / ("lib" + e)).read_bytes(), e
    assert (tmp_path / "out.dat").read_text() == (tmp_path / "lib.dat").read_text()
    assert (res.type > 0).mean() > 0.9

    if prog != "strack":
        return
    # cullst on the strack products against st_cull_run on the same grids
    run = subprocess.run([exe["cullst"], "-boxSize", "5", "-islandThresh", "4", str(tmp_path / "out"),
                          str(tmp_path / "culled")], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    print(run.stderr)
    assert run.returncode == 0, run.stderr
    hdr, g1, g2 = files.read_dat(str(tmp_path / "out.dat"))
    offR, offA, corr, typ = files.read_offsets(str(tmp_path / "out"), hdr[2], hdr[3])
    cp, arrs = track.cull_params(offR, offA, corr, typ, bR=5, bA=5, islandThresh=4)
    device.cull("st_cull_run", cp)
    shutil.copy(tmp_path / "out.dat", tmp_path / "lib2.dat")
    cm = files.CullMeta(hdr[0], hdr[1], hdr[4], hdr[5], 2, 8, 5, 17, -1, 0.75, 1.0, 0.0, None, None)
    files.write_cull(str(tmp_path / "libc"), str(tmp_path / "lib2"), cm, arrs["offRS"], arrs["offAS"],
                     arrs["sigmaR"], arrs["sigmaA"], arrs["corr"], arrs["type"])
    for e in RASTERS:
        assert (tmp_path / ("culled" + e)).read_bytes() == (tmp_path / ("libc" + e)).read_bytes(), e
    for e in (".sr", ".sa"):
        assert (tmp_path / ("out" + e)).read_bytes() == (tmp_path / ("lib2" + e)).read_bytes(), e
    assert abi.ST_OK == 0
