"""CPU checks of the C-ABI boundary: the library builds, loads and exports
every entry point include/*.h declares (no GPU call is made)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "big-int_amd", "libbigint.so")


def declared():
    names = []
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names += re.findall(r"\b(bi_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared()
    for must in ("bi_create", "bi_destroy", "bi_add_u32", "bi_add_u64", "bi_mul_u32", "bi_mul_u64",
                 "bi_add_u32_dev", "bi_mul_u32_dev"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libbigint.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libbigint.so not built")
def test_ffi_binds_every_declared_symbol():
    import bigint_ffi
    missing = [n for n in declared() if n not in bigint_ffi.SIGNATURES]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libbigint.so not built")
def test_error_strings_without_gpu():
    import bigint_ffi
    L = bigint_ffi.lib()
    assert L.bi_strerror(0) == b"ok"
    assert L.bi_strerror(-5) == b"division by zero"
    # argument validation happens before any device work
    assert L.bi_add_u32(None, 4, 1, None, None, None) == -1
    assert L.bi_mul_u32_dev(None, 4, 1, None, None, None) == -1
