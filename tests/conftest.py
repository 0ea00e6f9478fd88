import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def design_c1():
    """Benchmark configuration 0 (10 layers, 200x200, 20k nets), built."""
    from instantgr_amd.host import Design
    return Design.preset(0).build(8)


@pytest.fixture(scope="session")
def small_design():
    """A small design for quick checks (6 layers, 48x40, 300 nets)."""
    from instantgr_amd.host import Design
    return Design.custom(layers=6, xsize=48, ysize=40, nets=300, pin_cap=12, sigma=5.0, seed=7).build(4)
