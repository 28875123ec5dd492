import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device 0)")
    # build what is missing (a machine without the toolchain runs on the prebuilt files)
    lib = os.path.join(ROOT, "snusolver_amd", "libsnusolver.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-C", os.path.join(ROOT, "snusolver_amd", "csrc"), "-j8"], check=True)
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True)


@pytest.fixture(scope="session")
def S():
    import snusolver_amd
    return snusolver_amd


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    return oracle
