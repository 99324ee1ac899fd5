import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (ROOT, HERE):
    if p not in sys.path:
        sys.path.insert(0, p)

# Under pytest-xdist each worker loads its own OpenMP runtime for the oracle.
# A team of every CPU per worker makes the teams spin against each other: an
# oracle test that takes 0.1 s alone then takes minutes.  Each worker gets its
# share of the CPUs instead (read when the oracle library is first loaded).
_workers = int(os.environ.get("PYTEST_XDIST_WORKER_COUNT", "0") or 0)
if _workers > 1:
    os.environ["OMP_NUM_THREADS"] = str(max(1, (os.cpu_count() or 1) // _workers))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def golden():
    import numpy as np

    def load(name):
        return np.load(os.path.join(HERE, "golden", name + ".npz"))
    return load
