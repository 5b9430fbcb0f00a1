import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))

import __graft_entry__  # noqa: E402

__graft_entry__.load_package()

ORACLE_LIB = ROOT / "oracle" / "liboracle_abi41.so"
HIP_LIB = ROOT / "ssbl-cuda-xpbd_amd" / "libssbl_xpbd_hip_abi41.so"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def nb():
    from ssbl_xpbd_amd import native_backend
    return native_backend


@pytest.fixture(scope="session")
def scenes():
    from ssbl_xpbd_amd import scenes as s
    return s


@pytest.fixture(scope="session")
def oracle(nb):
    if not ORACLE_LIB.exists():
        pytest.fail(f"oracle library missing: {ORACLE_LIB} (run __graft_entry__.build())")
    return nb.load_library(str(ORACLE_LIB))


@pytest.fixture(scope="session")
def hip(nb):
    # GPU tests must run the HIP library; a missing build is a failure, not a skip.
    return nb.load_library(str(HIP_LIB))
