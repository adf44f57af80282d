import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle.Oracle()


@pytest.fixture(scope="session")
def ref_lib():
    import oracle
    if not oracle.reference_available():
        pytest.skip("oracle/_ref not built (reference tree absent)")
    return oracle.Reference()


@pytest.fixture(scope="session")
def engine():
    from tcodec_amd import me
    eng = me.Engine(0)
    yield eng
    eng.close()
