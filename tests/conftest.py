import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
ORACLE = os.path.join(ROOT, "oracle")
if ORACLE not in sys.path:
    sys.path.insert(0, ORACLE)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN


@pytest.fixture(scope="session")
def hip_ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import arma_amd
    ctx = arma_amd.Context(0)
    yield ctx
    ctx.close()


@pytest.fixture(scope="session")
def p521_params():
    import arma_amd
    return arma_amd.dcmt_create_many(521, 512, 20240611)
