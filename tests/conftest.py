import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle as O
    O.lib()
    return O


def load_golden_ggft(N):
    """GGFT golden vectors of the reference solver at resolution N (16, 21, 23),
    written by tests/golden/make_ggft_golden.py."""
    import numpy as np
    return dict(np.load(os.path.join(ROOT, "tests", "golden", f"ggft_n{N}.npz")))


@pytest.fixture(scope="session")
def golden_ggft():
    return load_golden_ggft(23)


@pytest.fixture(scope="session", params=[16, 21, 23], ids=lambda n: f"n{n}")
def golden_ggft_n(request):
    return load_golden_ggft(request.param)
