import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


def _oracle_path(name):
    path = os.path.join(ROOT, "oracle", name)
    if not os.path.exists(path):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    return path


@pytest.fixture(scope="session")
def oracle64():
    from lbm_porouschem_amd import _abi
    return _abi.CLib(_oracle_path("liblbm_oracle.so"))


@pytest.fixture(scope="session")
def oracle32():
    from lbm_porouschem_amd import _abi
    return _abi.CLib(_oracle_path("liblbm_oracle_f32.so"))


@pytest.fixture(scope="session")
def hip():
    """The product library; GPU tests fail loudly if it is missing or no device is visible."""
    from lbm_porouschem_amd import _abi
    return _abi.load_library()
