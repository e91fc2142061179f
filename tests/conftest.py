"""Shared pytest setup: the ``gpu`` marker, import paths, and the oracle/product handles."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
for p in (ROOT, ORACLE_DIR):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


def _ensure_oracle():
    try:
        import pomerol_oracle  # noqa: F401
    except ImportError:
        subprocess.run(["make", "-C", ORACLE_DIR], check=True, capture_output=True)
    import pomerol_oracle

    return pomerol_oracle


@pytest.fixture(scope="session")
def oracle():
    """The CPU oracle (test infrastructure only)."""
    return _ensure_oracle()


@pytest.fixture(scope="session")
def pa():
    """The product package; GPU tests fail (not skip) when the device or library is missing."""
    import pomerol_amd

    assert pomerol_amd.device_ok(), "no usable gfx950 device for a gpu-marked test"
    return pomerol_amd
