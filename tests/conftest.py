import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _ensure_built():
    """Build the engine and the oracle only when nothing is built yet."""
    from opentm_amd import build
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        build.build_oracle()
    if not os.path.exists(build.LIB):
        build.build_engine()
    import glob
    if not glob.glob(os.path.join(ROOT, "opentm_amd", "openTM*.so")):
        build.build_pybind()


_ensure_built()


@pytest.fixture(scope="session")
def orc():
    import oracle.oracle as o
    return o
