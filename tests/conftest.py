import os
import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (runs the HIP library)")


@pytest.fixture(scope="session")
def RBF():
    import rbfx_loader

    return rbfx_loader.load()


@pytest.fixture(scope="session")
def O():
    import oracle

    return oracle


@pytest.fixture(scope="session")
def gpu(RBF):
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    dev = RBF._lib.device_for("cuda:0")
    return dev
