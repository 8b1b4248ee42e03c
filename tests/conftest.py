import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)

REFERENCE_DIR = "/root/reference"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libindigo_hip.so)")
    config.addinivalue_line("markers", "slow: larger CPU-side checks")


def _ensure_built():
    """Build the native libraries only if they are missing (build() always rebuilds)."""
    from indigo2suite_amd import build
    if not os.path.exists(build.GRAPH_LIB):
        build.build_graph_lib()
    if not os.path.exists(build.ORACLE_LIB):
        build.build_oracle(with_reference=False)


_ensure_built()


@pytest.fixture(scope="session")
def gpu():
    """Initialise the HIP library once; GPU tests fail loudly if it cannot load."""
    from indigo2suite_amd import device
    device.init(0)
    return device


def reference_present():
    return os.path.isdir(REFERENCE_DIR)
