import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


@pytest.fixture(scope="session")
def pdpkg():
    return importlib.import_module("2d_pd_gpu_amd")


def say(msg):
    """Progress line for GPU runs (flushed so a killed run still names its case)."""
    print(msg, flush=True)
