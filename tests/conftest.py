"""Test-session setup: import paths, the ``gpu`` marker, shared fixtures.

Tests marked ``gpu`` need a visible HIP device and the built libmfhip.so;
the rest run on any host.  The CPU checker library (oracle/libmfcpu.so) is
built on first use if it is missing.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and libmfhip.so")


@pytest.fixture(scope="session")
def mfcpu():
    """ctypes handle of the single-thread FP32 CPU restatement."""
    from oracle import cpu
    return cpu.load()


@pytest.fixture(scope="session")
def gpu_device():
    """Index of the HIP device the GPU tests run on; fails loudly if absent."""
    from apogee_amd import _hip
    n = _hip.device_count()
    assert n > 0, "no HIP device visible: GPU tests need an MI355X"
    return 0
