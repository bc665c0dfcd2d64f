import ctypes as C
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libcorrfunc_ref.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def oracle():
    """The CPU oracle (FP64 restatement of the reference), bound with prefix oracle_."""
    from corrfunc_amd import _abi
    if not os.path.isfile(ORACLE_SO):
        import subprocess
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "liboracle.so"])
    return _abi.bind(C.CDLL(ORACLE_SO), "oracle_")


@pytest.fixture(scope="session")
def reference():
    """The reference library compiled from /root/reference (only where it was built)."""
    from corrfunc_amd import _abi
    if not os.path.isfile(REF_SO):
        pytest.skip("compiled reference not available (oracle/_ref not built)")
    return _abi.bind(C.CDLL(REF_SO))


@pytest.fixture(scope="session")
def engine():
    """The HIP engine; fails loudly if the library or the GPU is missing."""
    from corrfunc_amd import _lib
    return _lib.lib()
