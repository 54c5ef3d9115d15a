import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    # torch is imported before any test loads libvkvg.so (vkvg_amd.lib() does the same), so the
    # library resolves its HIP runtime to torch's copy and the process holds one runtime; torch's
    # GPU context is created up front. Counting devices does not initialise one.
    try:
        import torch
    except ImportError:
        return
    if torch.cuda.device_count() > 0:
        torch.zeros(1, device="cuda")
        torch.cuda.synchronize()


@pytest.fixture(scope="session")
def device():
    # torch creates its GPU context first, whatever order the tests run in: some GPU tests use
    # torch tensors beside the library
    try:
        import torch
        if torch.cuda.is_available():
            torch.zeros(1, device="cuda")
            torch.cuda.synchronize()
    except ImportError:
        pass
    import vkvg_amd
    dev = vkvg_amd.Device()
    yield dev
    dev.close()
