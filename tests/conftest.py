import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "parallel-vertex-cover-algorithms-on-gpus_amd")
for p in (ROOT, PKG, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libvcgpu.so on cuda:0)")


@pytest.fixture(scope="session")
def ctx():
    import vcgpu
    c = vcgpu.Context(0)
    yield c
    c.close()
