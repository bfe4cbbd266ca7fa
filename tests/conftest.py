import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def gpu():
    """Skip unless a GPU is present; the HIP library must then load (no fallback)."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from zodipy_amd import _lib

    _lib.load()
    return torch.device("cuda")


def pytest_report_header(config):
    return f"zodipy_amd root: {ROOT}; HIP lib: {os.environ.get('ZODI_HIP_LIB', 'in-tree')}"
