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


def _zero_fill(lines, i, src):
    """True when the run of copies from a<src> that holds line i starts right after
    `v_accvgpr_write_b32 a<src>, <immediate>`: a constant fill (accumulators set to zero on
    a path that skips their products), not a copy of a live sum."""
    j = i
    while j > 0 and re.match(r"v_accvgpr_mov_b32 a\d+, a%d$" % src, lines[j - 1]):
        j -= 1
    return j > 0 and re.match(r"v_accvgpr_write_b32 a%d, (-?\d+|0x[0-9a-fA-F]+)$" % src,
                              lines[j - 1]) is not None


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
            if len(src) == 1 and dst == list(range(dst[0], dst[0] + 4)) \
                    and not _zero_fill(lines, i, src.pop()):
                hits.append(fn)
    return hits


def test_pattern_detector():
    bad = "_Zk:\n v_accvgpr_mov_b32 a8, a14\n v_accvgpr_mov_b32 a9, a14\n" \
          " v_accvgpr_mov_b32 a10, a14\n v_accvgpr_mov_b32 a11, a14\n"
    good = "_Zk:\n v_accvgpr_mov_b32 a8, a14\n v_accvgpr_mov_b32 a9, a15\n" \
           " v_accvgpr_mov_b32 a10, a16\n v_accvgpr_mov_b32 a11, a17\n"
    fill = "_Zk:\n v_accvgpr_write_b32 a12, 0\n v_accvgpr_mov_b32 a13, a12\n" \
           " v_accvgpr_mov_b32 a14, a12\n v_accvgpr_mov_b32 a15, a12\n v_accvgpr_mov_b32 a8, a12\n" \
           " v_accvgpr_mov_b32 a9, a12\n v_accvgpr_mov_b32 a10, a12\n v_accvgpr_mov_b32 a11, a12\n"
    assert splat_copies(bad) == ["_Zk"] and splat_copies(good) == []
    assert splat_copies(fill) == []
    # a live accumulator copied after a constant write to another register is still caught
    assert splat_copies("_Zk:\n v_accvgpr_write_b32 a20, 0\n v_accvgpr_mov_b32 a8, a14\n"
                        " v_accvgpr_mov_b32 a9, a14\n v_accvgpr_mov_b32 a10, a14\n"
                        " v_accvgpr_mov_b32 a11, a14\n") == ["_Zk"]


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
