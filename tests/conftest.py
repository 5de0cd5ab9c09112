"""Shared fixtures.  GPU tests carry @pytest.mark.gpu; everything else runs on CPU."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "cuda-k-svd-denoising_amd")
ORACLE = os.path.join(ROOT, "oracle")
for p in (PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the device)")


def _ensure_built():
    """Build the library and the oracle only if nothing is built yet."""
    if not os.path.exists(os.path.join(PKG, "libksvd_hip.so")):
        subprocess.check_call(["make", "-s", "-C", PKG, "-j8"])
    if not os.path.exists(os.path.join(ORACLE, "build", "libksvd_oracle.so")):
        subprocess.check_call(["make", "-s", "-C", ORACLE])


@pytest.fixture(scope="session")
def orc():
    _ensure_built()
    import pyoracle
    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def kl():
    _ensure_built()
    import pyksvd
    pyksvd.lib()
    return pyksvd
