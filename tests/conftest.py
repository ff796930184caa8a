import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libbia_hip.so on the device)")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def ctx():
    from bia_3d_gpu_amd import Context
    c = Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def fixture1332():
    """The reference's own N=1332 run (OUT/ of the reference repo): IC, BIA deltas, DIM result."""
    from bia_3d_gpu_amd import io
    st = io.read_bbb(os.path.join(GOLDEN, "0001332_10^0.bbb"))[0]
    (dt, bia), = io.read_timed(os.path.join(GOLDEN, "0001332_10^0.BIA"))
    (t, dim), = io.read_timed(os.path.join(GOLDEN, "0001332_10^0.DIM"))
    return dict(state=st, bia=bia, dim=dim, dt=dt, t=t, eps=-6e-7, eps_dim=-3e-7)
