import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")

if os.path.join(ROOT, "oracle") not in sys.path:
    sys.path.insert(0, os.path.join(ROOT, "oracle"))


def load_package():
    """Import the host package from the ``ipeps.jl_amd`` directory as ``ipeps_amd``."""
    if "ipeps_amd" in sys.modules:
        return sys.modules["ipeps_amd"]
    pkg_dir = os.path.join(ROOT, "ipeps.jl_amd")
    spec = importlib.util.spec_from_file_location(
        "ipeps_amd", os.path.join(pkg_dir, "__init__.py"),
        submodule_search_locations=[pkg_dir])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["ipeps_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def pkg():
    return load_package()
