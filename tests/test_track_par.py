"""The strack parameter file (st_read_track_par: parseTrack,
Strack/parseTrack.c:13-190).  Host code, no GPU.

Pinned by parseTrack.c itself: its keywords, the %i formats of
parseValues (sscanf, so hex and octal forms parse as C does), which
keywords are required and the messages it stops with, the defaults, the
division by scalefactor and its quirks (an unreadable edgeA resets the
range pad; offsetmaskfile without offsetmaskdat leaves no geodat; a
baseline needs intgeodat; -noComplex clears baseFlag).  The RDF line
rules come from GrIMP's rdfSource, which is not in the reference tree:
case-insensitive keywords, first occurrence wins and "(units)" / "!"
comments are parity unpinned."""
import pytest

from specklesource_amd import files

FULL = """! strack parameters
image1 (-) = {d}/a.slc
image1par = {d}/a.par
Image2 = {d}/b.slc          ! second image
image2par (-) = {d}/b.par
initialoffsetfile = {d}/offsets.poly
outputfile = {d}/out/track
rstart (pixels) = 101
astart = 200
deltar = 24
deltaa = 0x18
nR = 10
nA = 012
wr = 64
wa = 64
wra = 32
waa = 32
edgepad = 5
scalefactor = 2
navgR = 2
navgA = 3
"""


def _write(tmp_path, text, name="track.par"):
    p = tmp_path / name
    p.write_text(text.format(d=tmp_path))
    return str(p)


def test_full_file(tmp_path):
    d = files.read_track_par(_write(tmp_path, FULL))
    assert d["image1"] == "%s/a.slc" % tmp_path and d["image2"] == "%s/b.slc" % tmp_path
    assert d["par1"] == "%s/a.par" % tmp_path and d["vrt1"] == "" and d["vrt2"] == ""
    assert d["polyShift"] == 1 and d["initialOffsetFile"].endswith("offsets.poly") and d["initShift"] == ""
    assert d["outputFile"] == "%s/out/track" % tmp_path
    # C integer division by scalefactor; %i reads 0x18 as 24 and 012 as 10
    assert (d["rStart"], d["aStart"], d["deltaR"], d["deltaA"]) == (50, 100, 12, 12)
    assert (d["nR"], d["nA"], d["wR"], d["wA"], d["wRa"], d["wAa"]) == (10, 10, 64, 64, 32, 32)
    assert (d["edgePad"], d["edgePadR"], d["edgePadA"], d["scaleFactor"]) == (5, 5, 5, 2)
    assert (d["navgR"], d["navgA"]) == (2, 3)
    assert d["maskFlag"] == 0 and d["intFlag"] == 0 and d["baseFlag"] == 0
    p = files.track_params_from_par(d, noComplex=1)
    assert (p.rStart, p.deltaA, p.nA, p.wRa, p.edgePadR, p.scaleFactor, p.polyShift, p.noComplex) == \
        (50, 12, 10, 32, 5, 2, 1, 1)


def test_first_occurrence_and_case(tmp_path):
    d = files.read_track_par(_write(tmp_path, FULL + "WR = 128\nrstart = 7\n"))
    assert d["wR"] == 64 and d["rStart"] == 50


@pytest.mark.parametrize("drop,msg", [
    ("image1 (-)", "Missing image1"),
    ("image1par", "Missing image1par and no vrt"),
    ("Image2", "Missing image2"),
    ("image2par", "Missing image2par and no vrt"),
    ("initialoffsetfile", "Missing initialoffsetfile"),
    ("outputfile", "missing ouput file"),
    ("rstart", "rstart not specified"),
    ("navgA", "navgA not specified"),
])
def test_required_keywords(tmp_path, drop, msg):
    text = "\n".join(l for l in FULL.split("\n") if not l.startswith(drop + " ") and not l.startswith(drop + "="))
    with pytest.raises(files.IoError, match=msg):
        files.read_track_par(_write(tmp_path, text))


def test_invalid_value(tmp_path):
    with pytest.raises(files.IoError, match="Invalid wr"):
        files.read_track_par(_write(tmp_path, FULL.replace("wr = 64", "wr = sixty-four")))


def test_vrt_replaces_par_file(tmp_path):
    (tmp_path / "a.slc.vrt").write_text("<VRTDataset/>")
    text = "\n".join(l for l in FULL.split("\n") if not l.startswith("image1par"))
    d = files.read_track_par(_write(tmp_path, text))
    assert d["vrt1"] == "%s/a.slc.vrt" % tmp_path and d["par1"] == "" and d["par2"].endswith("b.par")


def test_initial_offset_grids(tmp_path):
    d = files.read_track_par(_write(tmp_path, FULL + "initShift = {d}/init\n"))
    assert d["polyShift"] == 0 and d["initShift"] == "%s/init" % tmp_path


def test_edge_pads_and_quirk(tmp_path):
    d = files.read_track_par(_write(tmp_path, FULL + "edgeR = 7\nedgeA = 9\n"))
    assert (d["edgePadR"], d["edgePadA"]) == (7, 9)
    # an unreadable edgeA resets the range pad to edgepad and keeps edgePadA
    d = files.read_track_par(_write(tmp_path, FULL + "edgeR = 7\nedgeA = x\n"))
    assert (d["edgePad"], d["edgePadR"], d["edgePadA"]) == (5, 5, 5)
    # no edgepad: 0; unreadable scalefactor: 1
    text = FULL.replace("edgepad = 5\n", "").replace("scalefactor = 2", "scalefactor = two")
    d = files.read_track_par(_write(tmp_path, text))
    assert (d["edgePad"], d["scaleFactor"], d["rStart"]) == (0, 1, 101)


def test_masks(tmp_path):
    d = files.read_track_par(_write(tmp_path, FULL + "maskfile = {d}/m\nmaskgeodat = {d}/m.geodat\n"))
    assert (d["maskFlag"], d["maskType"]) == (1, 0) and d["maskGeodat"].endswith("m.geodat")
    with pytest.raises(files.IoError, match="Missing maskgeodat"):
        files.read_track_par(_write(tmp_path, FULL + "maskfile = {d}/m\n"))
    d = files.read_track_par(_write(tmp_path, FULL + "offsetmaskfile = {d}/om\noffsetmaskvrt = {d}/om.vrt\n"))
    assert (d["maskFlag"], d["maskType"]) == (1, 1) and d["maskFile"].endswith("om")
    assert d["maskGeodat"] == "" and d["maskVrt"].endswith("om.vrt")


def test_interferogram_and_baseline(tmp_path):
    d = files.read_track_par(_write(tmp_path, FULL + "intfile = {d}/i.int\nintgeodat = {d}/i.geodat\n"))
    assert d["intFlag"] == 1 and d["intFile"].endswith("i.int") and d["baseFlag"] == 0
    with pytest.raises(files.IoError, match="Missing intgeodat"):
        files.read_track_par(_write(tmp_path, FULL + "intfile = {d}/i.int\n"))
    with pytest.raises(files.IoError, match="Missing intgeodat"):
        files.read_track_par(_write(tmp_path, FULL + "baseparams = {d}/base\n"))
    text = FULL + "baseparams = {d}/base\nintgeodat = {d}/i.geodat\n"
    assert files.read_track_par(_write(tmp_path, text))["baseFlag"] == 1
    assert files.read_track_par(_write(tmp_path, text), no_complex=True)["baseFlag"] == 0


def test_missing_file(tmp_path):
    with pytest.raises(files.IoError):
        files.read_track_par(str(tmp_path / "none.par"))
