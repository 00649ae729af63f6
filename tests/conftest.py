import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "gbd-pcg_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def gpu():
    """Skips when no HIP device is visible; fails loudly when the library is missing."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import gbdpcg
    gbdpcg.lib()  # raises if libgbdpcg.so was not built
    return torch.device("cuda:0")
