import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: longer CPU checks")


@pytest.fixture(scope="session")
def gpu_ctx():
    import slr_amd
    return slr_amd.Context(0)


@pytest.fixture(scope="session")
def host_ctx():
    import slr_amd
    return slr_amd.Context(-1)


def need_ref():
    from oracle import refdrv
    if not refdrv.available():
        pytest.skip("reference harness not built (make -f oracle/ref.mk)")
    return refdrv
