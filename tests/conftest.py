"""Shared fixtures.  GPU tests are marked `gpu`; everything else runs on the CPU."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import __graft_entry__ as _entry  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def gc():
    """The gridcalc_amd package (libgridcalc_hip.so must be built)."""
    return _entry.load_package()


@pytest.fixture(scope="session")
def orc():
    """The CPU oracle."""
    mod = _entry.load_oracle()
    mod.lib()
    return mod


@pytest.fixture(scope="session")
def ctx(gc):
    c = gc.Context()
    yield c
    c.close()
