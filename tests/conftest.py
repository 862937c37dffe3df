import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) device")


@pytest.fixture(scope="session")
def ctx():
    # torch's HIP runtime must initialise the device before the library's
    # context does when one process uses both (torch.cuda reports no GPU
    # otherwise); tests hand torch device buffers to the library
    import torch
    if torch.cuda.device_count() > 0:
        torch.cuda.init()
    from jetshow_amd import host
    c = host.Context(0)
    yield c
    c.close()
