import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "ar-physics_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle as O

    if not os.path.exists(O.LIB):
        O.build()
    return O


def gpu_available():
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False
