import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (run with -m gpu)")


@pytest.fixture(scope="session")
def ref_lib():
    """The compiled reference (oracle/_ref); built here when the sources exist."""
    from oracle import ref
    try:
        ref.build()
    except Exception as e:  # pragma: no cover - toolchain problems
        pytest.skip(f"reference build failed: {e}")
    if not ref.available():
        pytest.skip("oracle/_ref/libmufft_ref.so is not available")
    return ref


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import mufft_amd
    if mufft_amd.device_count() < 1:
        pytest.skip("no HIP device")
    return torch.device("cuda:0")
