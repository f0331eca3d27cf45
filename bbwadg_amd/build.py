"""Build libbbwadg_hip.so in-tree with hipcc for gfx950.

    python -m bbwadg_amd.build          # or __graft_entry__.build()
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libbbwadg_hip.so")
SOURCES = ["bbw_abi.hip", "bbw_acoustic.hip", "bbw_fused.hip", "bbw_elastic.hip", "bbw_adjoint.hip", "bbw_batch.hip", "bbw_energy.hip",
           "bbw_fused_elas.hip", "bbw_slice.hip", "bbw_curved.hip", "bbw_adjoint_elas.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
    # no packed FP32 instructions in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def sources():
    return [os.path.join(CSRC, s) for s in SOURCES]


def build(verbose=True):
    cmd = [HIPCC] + FLAGS + ["-I" + os.path.join(ROOT, "include"), "-o", LIB + ".tmp"] + sources()
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build()
