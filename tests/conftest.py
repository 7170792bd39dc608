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


def film_mismatch(gpu, ref):
    """Pixels whose bits differ anywhere in their three channels, and the
    largest relative difference.  With the shared counter RNG the library
    and the oracle trace the same paths with the same operation order; sin
    and cos of sampling angles are evaluated in fp64 and rounded once on
    both sides, which can still differ in the last fp32 bit where the two
    fp64 results straddle a rounding boundary (about one angle in 2^29),
    so a bound of a pixel or two is kept for that case."""
    import numpy as np
    g = np.ascontiguousarray(gpu, np.float32)
    r = np.ascontiguousarray(ref, np.float32)
    diff = (g.view(np.uint32) != r.view(np.uint32)).reshape(-1, g.shape[-1]).any(axis=1)
    rel = np.abs(g - r) / np.maximum(np.abs(r), 1e-3)
    return int(diff.sum()), float(rel.max()) if rel.size else 0.0


def outlier_bound(n, alpha=1e-4, p=0.0027, dof=None):
    """Largest number of |z| > 3 entries that n correct z-scores exceed with
    probability below alpha: the binomial quantile around p = 0.27 % for
    Gaussian z, or, when a variance comes from `dof`-degree-of-freedom
    replicates, around P(|t_dof| > 3)."""
    from scipy import stats
    if dof is not None:
        p = float(2.0 * stats.t.sf(3.0, dof))
    return int(stats.binom.ppf(1.0 - alpha, n, p))
