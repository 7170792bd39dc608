import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs through the HIP library)")
    config.addinivalue_line("markers", "slow: full-size workload")


def load_ar():
    """Import the agent-ray_amd package (its directory name has a hyphen)."""
    if "agent_ray_amd" in sys.modules:
        return sys.modules["agent_ray_amd"]
    pkg = os.path.join(ROOT, "agent-ray_amd")
    spec = importlib.util.spec_from_file_location("agent_ray_amd", os.path.join(pkg, "__init__.py"),
                                                  submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["agent_ray_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def ar():
    return load_ar()


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle


@pytest.fixture(scope="session")
def wl(ar):
    return ar.workloads
