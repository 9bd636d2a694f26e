import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (ROCm GPU); run with -m gpu")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.set_threads(min(8, os.cpu_count() or 1))
    return O


@pytest.fixture(scope="session")
def isa():
    import imagestackalignator_amd as isa

    isa.lib()
    return isa


@pytest.fixture(scope="session")
def gpu():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(autouse=True)
def _release_device_tensors():
    yield
    import sys as _sys

    mod = _sys.modules.get("tests.gpu_util")
    if mod is not None:
        import torch

        if torch.cuda.is_available():
            torch.cuda.synchronize()
        mod.KEEP.clear()
