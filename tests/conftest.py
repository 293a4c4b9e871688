"""Shared test setup: repo root on sys.path and the `gpu` marker.

Tests marked `gpu` need an MI355X and the built libxpbd_hip.so; everything
else runs on the CPU (oracle, host topology code, ABI surface).
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X) and the built libxpbd_hip.so")
    config.addinivalue_line("markers", "slow: long CPU test (large meshes)")


@pytest.fixture(scope="session")
def O():
    """The CPU oracle (test infrastructure)."""
    from oracle import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def X():
    """The product package (C ABI over libxpbd_hip.so)."""
    import bevy_xpbd_amd
    bevy_xpbd_amd.lib()
    return bevy_xpbd_amd
