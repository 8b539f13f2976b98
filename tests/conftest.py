import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("", "oracle", os.path.join("tests", "hostemu")):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)

# the oracle runs its loops under OpenMP with one thread per CPU by default; parallel test workers
# (pytest-xdist) would each start a full team and oversubscribe the CPUs many times over
if os.environ.get("PYTEST_XDIST_WORKER") and "OMP_NUM_THREADS" not in os.environ:
    os.environ["OMP_NUM_THREADS"] = "1"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libhcamd.so")


def rel_err(a, b):
    """Normwise relative error max|a-b| / max|b| (per the bench metric)."""
    import numpy as np
    d = np.max(np.abs(np.asarray(a) - np.asarray(b)))
    s = np.max(np.abs(np.asarray(b)))
    return float(d / s) if s > 0 else float(d)


@pytest.fixture(scope="session")
def gpu_ctx_factory():
    import hyperboliconvergence_amd as hca

    def make():
        return hca.Context(0)
    return make
