import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def small_data():
    import busseq_gpu_amd as bq
    n_b = [40, 30, 30]
    G, K = 200, 3
    Y, W, X = bq.simulate(n_b, G, K, 1001, want_truth=True)
    return {"Y": Y, "n_b": n_b, "G": G, "K": K, "W": W, "X": X}
