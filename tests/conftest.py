import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "necsim-rust_amd")
ORACLE = os.path.join(ROOT, "oracle")
for p in (PKG, ORACLE, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the GPU)")


def _ensure_built():
    # The GPU box receives prebuilt libraries; building here only happens on a
    # CPU checkout where they are missing.
    lib = os.path.join(PKG, "libnecsim_hip.so")
    if not os.path.exists(lib) and os.path.exists("/opt/rocm/bin/hipcc"):
        subprocess.run([sys.executable, os.path.join(PKG, "build.py")], check=True)
    olib = os.path.join(ORACLE, "_build", "libnecsim_oracle.so")
    if not os.path.exists(olib):
        subprocess.run(["make", "-s", "-C", ORACLE], check=True)


_ensure_built()


@pytest.fixture(scope="session")
def ctx():
    import necsim_hip as nh
    return nh.Context.get(0)
