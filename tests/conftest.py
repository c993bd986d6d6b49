import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "gales-multi-fluid_amd")
ORACLE = os.path.join(ROOT, "oracle")
for p in (ORACLE, PKG, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    # One HIP runtime per process: torch ships its own libamdhip64.so.7.  When
    # torch loads first, the library's libamdhip64.so.7 dependency binds to
    # that same runtime; loaded the other way round, torch finds no device.
    # Runs against a mutated oracle copy (test_oracle_mutants.py) are CPU-only.
    if os.environ.get("GMC_ORACLE_LIB"):
        return
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.init()
    except ImportError:
        pass


@pytest.fixture(scope="session", autouse=True)
def _oracle_built():
    subprocess.run(["make", "-s", "-C", ORACLE], check=True)


@pytest.fixture(scope="session")
def shim():
    """The product's element headers compiled for the CPU (tests/host_shim)."""
    import ctypes as C
    out = os.path.join(ROOT, "tests", "host_shim", "libelement_host.so")
    src = os.path.join(ROOT, "tests", "host_shim", "element_host.cpp")
    hdrs = [os.path.join(PKG, "csrc", f) for f in os.listdir(os.path.join(PKG, "csrc")) if f.endswith(".hpp")]
    if not os.path.exists(out) or any(os.path.getmtime(h) > os.path.getmtime(out) for h in hdrs + [src]):
        subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wno-unknown-pragmas",
                        "-I", os.path.join(PKG, "csrc"), "-o", out, src], check=True)
    return C.CDLL(out)
