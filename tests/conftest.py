"""Shared fixtures: the library binding, the oracle (checker only) and the gpu marker."""
import importlib.util
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "oracle"))


def load_pkg():
    name = "mcstar_libs_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(REPO, "mcstar-libs_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI of libmcstar_hip.so)")


@pytest.fixture(scope="session")
def mcs():
    m = load_pkg()
    m.lib()
    return m


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle
