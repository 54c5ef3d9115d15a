import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    # torch (its own HIP runtime) creates its GPU context before any test loads libvkvg.so: loading
    # the library registers its kernels with the HIP runtime it links, and a later torch context
    # then leaves the library's device creation failing. Counting devices does not initialise one.
    import torch
    if torch.cuda.device_count() > 0:
        torch.zeros(1, device="cuda")
        torch.cuda.synchronize()


@pytest.fixture(scope="session")
def device():
    # torch (its own HIP runtime) creates its GPU context first, whatever order the tests run in:
    # some GPU tests use torch tensors beside the library
    import torch
    if torch.cuda.is_available():
        torch.zeros(1, device="cuda")
        torch.cuda.synchronize()
    import vkvg_amd
    dev = vkvg_amd.Device()
    yield dev
    dev.close()
