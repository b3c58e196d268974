import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built HIP library")


@pytest.fixture(scope="session")
def qf_ctx():
    import torch
    import quickfoil_amd
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a GPU")
    ctx = quickfoil_amd.Context(0)
    yield ctx
    ctx.close()
