"""Tracking from inputs read through the file readers (initial-offset
polynomial file, initial-offset grids, offset mask) gives bit-identical
results to the same inputs passed in memory, and the offsets written back
out (.dr/.da/.cc/.mt) read back unchanged."""
import numpy as np
import pytest

from specklesource_amd import abi, configs, files, track
from tests.helpers import small_params

pytestmark = pytest.mark.gpu


def _scene():
    s = configs._spec(400, 400, 23, dr0=2.4, da0=-1.7, gmin=0.7, gmax=0.7)
    return track.synth_host(s)


def test_track_from_files_equals_in_memory(device, tmp_path):
    a, b = _scene()
    # polynomial file
    (tmp_path / "init.off").write_text("range_offset_polynomial: 2.4 0 0\nazimuth_offset_polynomial: -1.7 0 0\n")
    r, az = files.read_offset_polys(str(tmp_path / "init.off"))
    p_file = small_params(32, 400, 400, step=16, rShiftPoly=r, aShiftPoly=az)
    p_mem = small_params(32, 400, 400, step=16, rShiftPoly=[2.4, 0, 0], aShiftPoly=[-1.7, 0, 0])
    # initial-offset grids with holes
    gr = np.full((10, 10), 2.4, np.float32)
    ga = np.full((10, 10), -1.7, np.float32)
    gr[3:5, 2:6] = ga[3:5, 2:6] = -2e9
    (tmp_path / "prior.da.dat").write_text("0 0 10 10 40.0 40.0\n")
    (tmp_path / "prior.da").write_bytes(ga.astype(">f4").tobytes())
    (tmp_path / "prior.dr").write_bytes(gr.astype(">f4").tobytes())
    grids = files.read_initial_offsets(str(tmp_path / "prior.da"), str(tmp_path / "prior.dr"))
    q_file = small_params(32, 400, 400, step=16)
    keep = files.apply_initial_offsets(q_file, grids)
    q_mem = small_params(32, 400, 400, step=16, polyShift=0)
    q_mem.initDr, q_mem.initDa = abi.ptr(gr, abi.C.c_float), abi.ptr(ga, abi.C.c_float)
    q_mem.initNr = q_mem.initNa = 10
    q_mem.initR0, q_mem.initA0, q_mem.initDeltaR, q_mem.initDeltaA = 0.0, 0.0, 40.0, 40.0
    # mask
    m = np.ones((20, 20), np.uint8)
    m[5:9, 3:12] = 0
    (tmp_path / "mask.dat").write_text("0 0 20 20 20.0 20.0\n")
    (tmp_path / "mask.bytes").write_bytes(m.tobytes())
    hdr, gm = files.read_mask(str(tmp_path / "mask.dat"), str(tmp_path / "mask.bytes"))
    files.apply_mask(p_file, hdr, gm)
    p_mem.mask = abi.ptr(m, abi.C.c_uint8)
    p_mem.maskNr = p_mem.maskNa = 20
    p_mem.maskNrl = p_mem.maskNal = 20
    pid = device.upload_pair(a, b)
    try:
        runs = [(device.speckle_track(pid, x), device.speckle_track(pid, y)) for x, y in ((p_file, p_mem),
                                                                                         (q_file, q_mem))]
    finally:
        device.release_pair(pid)
    for g1, g2 in runs:
        for f in ("offR", "offA", "corr", "type"):
            assert np.array_equal(getattr(g1, f), getattr(g2, f)), f
    assert (runs[0][0].type == 0).any()  # the mask removed points
    # write the products and read them back
    g = runs[1][0]
    meta = files.offsets_meta(q_file, "geodat.in", "/d/p2/slc.2", "/d/p1/slc.1")
    files.write_result(str(tmp_path / "out"), meta, g)
    r_, a_, c_, t_ = files.read_offsets(str(tmp_path / "out"), g.offR.shape[1], g.offR.shape[0])
    assert r_.tobytes() == g.offR.tobytes() and a_.tobytes() == g.offA.tobytes() and c_.tobytes() == g.corr.tobytes()
    del keep
