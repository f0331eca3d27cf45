"""Compiles tests/c_abi/c_caller.c, a plain-C program that includes include/bbwadg.h and links
libbbwadg_hip.so as a C caller of the reference's host code would, and runs it: the header
must compile as strict C99 without warnings and the argument checks must return the
documented BBW_ERR_* classes before any device work (no GPU needed)."""
import os
import shutil
import subprocess

import pytest

from bbwadg_amd import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not found")
def test_plain_c_caller(tmp_path):
    libdir = os.path.dirname(_lib.LIB_PATH)
    exe = str(tmp_path / "c_caller")
    cmd = ["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic",
           "-I" + os.path.join(ROOT, "include"), os.path.join(HERE, "c_abi", "c_caller.c"),
           "-L" + libdir, "-lbbwadg_hip", "-Wl,-rpath," + libdir, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all checks passed" in r.stdout
