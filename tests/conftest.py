import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("", "oracle", "big-int_amd"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libbigint.so")


@pytest.fixture(scope="session")
def ctx():
    import bigint_ffi

    c = bigint_ffi.Context(0)
    yield c
    c.close()
