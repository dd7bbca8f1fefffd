"""The C ABI library loads without a GPU, exports every symbol the public
headers declare, and its structure layouts match the ctypes mirror."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from specklesource_amd import abi, track

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("speckle_amd.h", "speckle_synth.h", "speckle_io.h")]


def declared_functions():
    names = []
    for h in HEADERS:
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names += re.findall(r"^\s*(?:int|void|const char \*)\s*\**\s*(st_\w+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


def test_library_exports_declared_symbols():
    lib = track.load_library()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n


def test_struct_layouts_match_header():
    src = r'''
#include <stdio.h>
#include "speckle_amd.h"
#include "speckle_synth.h"
#include "speckle_io.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(st_track_params), sizeof(st_track_out),
         sizeof(st_track_diag), sizeof(st_cull_params), sizeof(st_synth_spec), sizeof(st_job),
         sizeof(st_job_info), sizeof(st_offsets_meta), sizeof(st_cull_meta));
  printf("%zu %zu %zu %zu\n", sizeof(st_cull_ls_params), sizeof(st_cull_ls_meta), sizeof(st_track_par_file),
         sizeof(st_track_stats));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "t")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    sizes = [int(x) for x in out]
    from specklesource_amd import files
    assert sizes == [C.sizeof(abi.TrackParams), C.sizeof(abi.TrackOut), C.sizeof(abi.TrackDiag),
                     C.sizeof(abi.CullParams), C.sizeof(abi.SynthSpec), C.sizeof(abi.Job),
                     C.sizeof(abi.JobInfo), C.sizeof(files.OffsetsMeta), C.sizeof(files.CullMeta),
                     C.sizeof(abi.CullLSParams), C.sizeof(files.CullLSMeta), C.sizeof(files.TrackParFile),
                     C.sizeof(abi.TrackStats)]


def test_ctx_create_without_gpu_reports_nodev():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = track.load_library()
    ctx = C.c_void_p()
    assert lib.st_ctx_create(0, C.byref(ctx)) == abi.ST_ERR_NODEV


def test_server_create_without_gpu_reports_nodev():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = track.load_library()
    srv = C.c_void_p()
    assert lib.st_server_create(0, 2, 1 << 30, C.byref(srv)) == abi.ST_ERR_NODEV
    assert lib.st_server_create(0, 0, 1 << 30, C.byref(srv)) == abi.ST_ERR_ARG


def test_host_synth_is_deterministic():
    from specklesource_amd import configs
    s, _ = configs.cfg1(scale=16)
    a1, b1 = track.synth_host(s)
    a2, b2 = track.synth_host(s)
    assert (a1 == a2).all() and (b1 == b2).all()
    assert 0.5 < (abs(a1) ** 2).mean() < 2.5
