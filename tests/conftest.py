import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libepet.so)")
    config.addinivalue_line("markers", "slow: long-running CPU check")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def gpu_ctx_factory():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from epetreco_amd.api import Context

    def make(workload_name=None, setup=None, grid=None, nrays=8, seed=0):
        from epetreco_amd import geometry as G
        ctx = Context(0)
        if workload_name is not None:
            setup, grid, nrays = G.workload(workload_name)
        if setup is not None:
            ctx.set_geometry(setup, grid, nrays=nrays, seed=seed)
        return ctx

    return make
