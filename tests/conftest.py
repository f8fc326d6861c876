import ctypes
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def _gpu_count():
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        return 0
    n = ctypes.c_int(0)
    if hip.hipGetDeviceCount(ctypes.byref(n)) != 0:
        return 0
    return n.value


@pytest.fixture(scope="session")
def gpu():
    if _gpu_count() < 1:
        pytest.skip("no GPU visible")
    import channel_rl_amd
    channel_rl_amd.load_library()
    return True


@pytest.fixture(scope="session")
def wall_network(gpu):
    """The learned wall-stress closure trained as the benchmark trains it
    (10^6 samples of the small channel's runs, 20 epochs): (network, losses)."""
    from channel_rl_amd import wallnet
    X, Y = wallnet.solver_samples(1_000_000, seed=7)
    return wallnet.train(X, Y, epochs=20, batch=1024, lr=1e-3, seed=0)
