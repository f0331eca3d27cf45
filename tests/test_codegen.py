"""The AMD code of the MFMA kernels has no accumulator copy of the shape that made the exact
BB-WADG update's product loop wrong at N = 3: four consecutive v_accvgpr_mov_b32 into aX..aX+3
from one source AGPR (one component of an accumulator copied into all four of another) on a path
that skips an MFMA.  Compiles the device code to assembly with the build's flags (no GPU)."""
import os
import re
import subprocess
import tempfile

import pytest

from bbwadg_amd import build

SOURCES = ["bbw_fused.hip", "bbw_fused_elas.hip", "bbw_adjoint.hip", "bbw_batch.hip"]


def splat_copies(asm):
    lines = [l.strip() for l in asm.splitlines()]
    hits, fn = [], None
    for i in range(len(lines) - 3):
        if re.match(r"^_Z\w+:", lines[i]):
            fn = lines[i].split(":")[0]
        m = [re.match(r"v_accvgpr_mov_b32 a(\d+), a(\d+)$", lines[i + k]) for k in range(4)]
        if all(m):
            dst = [int(x.group(1)) for x in m]
            src = {int(x.group(2)) for x in m}
            if len(src) == 1 and dst == list(range(dst[0], dst[0] + 4)):
                hits.append(fn)
    return hits


def test_pattern_detector():
    bad = "_Zk:\n v_accvgpr_mov_b32 a8, a14\n v_accvgpr_mov_b32 a9, a14\n" \
          " v_accvgpr_mov_b32 a10, a14\n v_accvgpr_mov_b32 a11, a14\n"
    good = "_Zk:\n v_accvgpr_mov_b32 a8, a14\n v_accvgpr_mov_b32 a9, a15\n" \
           " v_accvgpr_mov_b32 a10, a16\n v_accvgpr_mov_b32 a11, a17\n"
    assert splat_copies(bad) == ["_Zk"] and splat_copies(good) == []


@pytest.mark.parametrize("src", SOURCES)
def test_no_splat_accumulator_copies(src):
    if not os.path.exists(build.HIPCC):
        pytest.skip("hipcc not installed")
    flags = [f for f in build.FLAGS if f not in ("-shared", "-fPIC")]
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "k.s")
        cmd = [build.HIPCC] + flags + ["--cuda-device-only", "-S", "-I" + os.path.join(build.ROOT, "include"),
                                       "-o", out, os.path.join(build.CSRC, src)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-2000:]
        hits = splat_copies(open(out).read())
    assert hits == [], hits
