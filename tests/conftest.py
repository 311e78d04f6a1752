import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(autouse=True)
def _oracle_math_for_device_tests(request):
    """GPU tests compare the device's FP64 paths with the oracle's bit for bit, so
    the oracle evaluates the path's exp/log as the device does (exp_log.h); every
    other test keeps glibc's, the reference's own."""
    import oracle as O
    prev = O.DEVICE_MATH
    O.DEVICE_MATH = request.node.get_closest_marker("gpu") is not None
    yield
    O.DEVICE_MATH = prev
