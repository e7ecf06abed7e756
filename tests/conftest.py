import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "sol-rs_amd", "python"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 (MI355X) device; runs the HIP library")


@pytest.fixture(scope="session")
def ctx():
    import solhip
    c = solhip.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def oracle():
    import solref
    solref.lib()
    return solref
