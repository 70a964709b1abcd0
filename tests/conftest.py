"""Shared fixtures.  GPU tests are marked `gpu`; everything else runs on the CPU."""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def load_pkg():
    """Import svdag-compression_amd/ (a directory name with a hyphen) as `svdag_amd`."""
    if "svdag_amd" in sys.modules:
        return sys.modules["svdag_amd"]
    d = os.path.join(ROOT, "svdag-compression_amd")
    spec = importlib.util.spec_from_file_location("svdag_amd", os.path.join(d, "__init__.py"),
                                                  submodule_search_locations=[d])
    m = importlib.util.module_from_spec(spec)
    sys.modules["svdag_amd"] = m
    spec.loader.exec_module(m)
    return m


load_pkg()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")


@pytest.fixture(scope="session")
def svd():
    return load_pkg()


@pytest.fixture(scope="session")
def ctx(svd):
    c = svd.Context(0)
    yield c
    c.close()
