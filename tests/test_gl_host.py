"""Host build of the prime-field arithmetic behind the transform products
(big-int_amd/csrc/bi_ntt.h, p = 2^64 - 2^32 + 1) checked against 128-bit
integer arithmetic: four million add/sub/mul pairs (random and edge values),
a million reductions of random 128-bit values, root-of-unity orders,
inverses, and a length-64 cyclic convolution by transform against the direct
sum.  CPU only."""
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_prime_field_host():
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "gl_check")
        subprocess.run(["g++", "-O2", "-I", os.path.join(ROOT, "big-int_amd", "csrc"), "-o", exe,
                        os.path.join(ROOT, "tests", "gl_check.cpp")], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    assert "bad 0" in out
