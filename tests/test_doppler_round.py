"""doppler_round (the acceptance test's Doppler rounding by sign tests
against roots of unity, st_track_common.h) against the atan2 rounding of
estDopCarrier1 (dopplerEstimation.c:42-88): random lag sums over 40 decades
of magnitude, signed zeros / axes / extreme magnitudes, and angle sums
1-3e-9 rad either side of every rounding boundary, for w = 16..128.  Built
for the host with hipcc (no GPU needed)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_doppler_round_matches_atan2_rounding(tmp_path):
    exe = str(tmp_path / "doppler_round_check")
    src = os.path.join(HERE, "native", "doppler_round_check.hip")
    inc = os.path.join(ROOT, "specklesource_amd", "csrc")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "--offload-arch=gfx950", "-I", inc, "-I",
                    os.path.join(ROOT, "include"), src, "-o", exe], check=True, capture_output=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    vals = dict(zip(out[0::3], zip(map(int, out[1::3]), map(int, out[2::3]))))
    for kind in ("random", "special", "near"):
        bad, n = vals[kind]
        assert n > 0 and bad == 0, (kind, bad, n)
