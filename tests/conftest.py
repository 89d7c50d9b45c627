import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "edge-flow-maps_amd")
for p in (ROOT, PKG, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built libedge_hip.so")


@pytest.fixture(scope="session")
def orc32():
    from oracle.oracle import Oracle
    import numpy as np
    return Oracle(np.float32)


@pytest.fixture(scope="session")
def orc64():
    from oracle.oracle import Oracle
    import numpy as np
    return Oracle(np.float64)


@pytest.fixture(scope="session")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    import edge_hip
    edge_hip.lib()  # raises if libedge_hip.so is missing: GPU tests must run the HIP path
    return torch.device("cuda:0")
