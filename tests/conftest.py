import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / "oracle", ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libntt_hip")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle():
    import ntt_oracle
    ntt_oracle.lib()
    return ntt_oracle


@pytest.fixture(scope="session")
def hip():
    """The HIP library path; the test fails loudly when the build is missing."""
    if not gpu_available():
        pytest.fail("GPU test collected without a HIP device")
    from entity_amd import _abi
    _abi.load(False)
    _abi.load(True)
    return _abi
