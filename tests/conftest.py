import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through liblagpm.so on the device)")


def gpu_available() -> bool:
    try:
        import torch  # noqa: F401  (device discovery only)
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def ctx():
    from lageunha_amd import lagpm
    c = lagpm.Context(0)
    yield c
    c.close()
