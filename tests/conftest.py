import os
import sys

import pytest

# torch bundles its own HIP runtime under the same soname as /opt/rocm's (libamdhip64.so.7), and
# the first one loaded serves the whole process.  GPU tests that use torch tensors need torch's,
# so torch is imported before anything loads libnq_hip.so (bench.py does the same).
try:
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libnq_hip.so on the device)")
    config.addinivalue_line("markers", "slow: CPU test that takes more than ~20 s")


@pytest.fixture(scope="session")
def published():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "published.json")) as f:
        d = json.load(f)
    d["solutions"] = {int(k): int(v) for k, v in d["solutions"].items()}
    return d


@pytest.fixture(scope="session")
def ref_vectors():
    import numpy as np
    return np.load(os.path.join(ROOT, "tests", "golden", "ref_vectors.npz"))
