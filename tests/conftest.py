import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on cuda:0)")
    config.addinivalue_line("markers", "slow: full-size workloads (still bounded to ~1 min)")


@pytest.fixture(scope="session")
def gpu_ctx():
    """One device context for the whole GPU session (fails loudly without a GPU)."""
    import fastapsp_amd as fa
    ctx = fa.Context(0)
    yield ctx
    ctx.close()
