import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "crystal-opt_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def ctx():
    """One lj_ctx on cuda:0 bound to torch's current stream."""
    import torch
    import lipjoin
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    c = lipjoin.Context(0)
    yield c
    c.close()
