import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X)")


@pytest.fixture(scope="session")
def orc64():
    from oracle import Oracle
    return Oracle("float64")


@pytest.fixture(scope="session")
def orc32():
    from oracle import Oracle
    return Oracle("float32")


@pytest.fixture(scope="session")
def ref():
    from oracle import Reference
    if not Reference.available():
        pytest.skip("oracle/_ref/libmpdata_ref.so not built (needs the reference tree)")
    return Reference()


@pytest.fixture(scope="session")
def ref32():
    import numpy as np
    from oracle import Reference
    if not Reference.available(np.float32):
        pytest.skip("oracle/_ref/libmpdata_ref32.so not built (needs the reference tree)")
    return Reference(np.float32)
