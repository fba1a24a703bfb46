import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built libsdpd.so")


def pytest_collection_modifyitems(config, items):
    # GPU tests run only where a GPU is present (select them with -m gpu)
    pass


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


def _device_normals(seed, lo, hi, step):
    from mcf_amd.sdpd import pair_normals

    return pair_normals(lo, hi, seed, step)


@pytest.fixture(autouse=True)
def device_noise_for_gpu_tests(request):
    """GPU tests compare the device with the oracle on the device's thermal
    noise: the oracle takes the normals of each evaluation's pairs from the
    device's stream (sdpd_pair_normals).  CPU tests keep the oracle's own
    stream."""
    if request.node.get_closest_marker("gpu") is None:
        yield
        return
    from oracle import oracle as O

    O.set_noise_source(_device_normals)
    try:
        yield
    finally:
        O.set_noise_source(None)
