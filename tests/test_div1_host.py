"""Host build of the one-digit division routines (big-int_amd/csrc/bi_div1.h,
Moller-Granlund division by an invariant digit) checked against plain 64-bit
and 128-bit integer division on 12 million random and edge cases.  CPU only."""
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_div1_helpers_host():
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "div1_check")
        subprocess.run(["g++", "-O2", "-I", os.path.join(ROOT, "big-int_amd", "csrc"), "-o", exe,
                        os.path.join(ROOT, "tests", "div1_check.cpp")], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    assert "bad 0" in out
