import glob
import hashlib
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")
CSRC = os.path.join(ROOT, "snusolver_amd", "csrc")
LIB = os.path.join(ROOT, "snusolver_amd", "libsnusolver.so")


def source_hash():
    """Hash of the library's sources, computed as its Makefile computes the
    build id baked into the library (same files, same order)."""
    names = sorted(os.path.basename(p) for ext in ("*.cpp", "*.hip", "*.h")
                   for p in glob.glob(os.path.join(CSRC, ext)))
    files = [os.path.join(CSRC, n) for n in names]
    files += [os.path.join(ROOT, "include", "snusolver.h"),
              os.path.join(ROOT, "include", "snusolver_compat.hpp")]
    h = hashlib.sha1()
    for f in files:
        with open(f, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def built_id():
    """Build id of the library on disk (read in a child process, so this
    process does not keep a stale library loaded)."""
    code = ("import ctypes,sys; L=ctypes.CDLL(sys.argv[1]); "
            "L.snu_build_id.restype=ctypes.c_char_p; print(L.snu_build_id().decode())")
    r = subprocess.run([sys.executable, "-c", code, LIB], capture_output=True, text=True)
    return r.stdout.strip() if r.returncode == 0 else None


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device 0)")
    # rebuild whenever the library is missing or was built from other sources
    # than the ones in the tree (a stale prebuilt library is never tested)
    if not os.path.exists(LIB) or built_id() != source_hash():
        subprocess.run(["make", "-C", CSRC, "clean"], check=True, stdout=subprocess.DEVNULL)
        subprocess.run(["make", "-C", CSRC, "-j8"], check=True)
        if built_id() != source_hash():
            raise RuntimeError("libsnusolver.so build id does not match the sources after a rebuild")
    ora = os.path.join(ROOT, "oracle", "liboracle.so")
    src = os.path.join(ROOT, "oracle", "snu_oracle.c")
    if not os.path.exists(ora) or os.path.getmtime(ora) < os.path.getmtime(src):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "-B"], check=True)


@pytest.fixture(scope="session")
def S():
    import snusolver_amd
    return snusolver_amd


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    return oracle
