import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "oracle"), os.path.join(ROOT, "zedm-x_amd"), os.path.join(ROOT, "tests", "golden"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs through the C-ABI library")


@pytest.fixture(scope="session")
def golden_geometry():
    import json

    with open(os.path.join(ROOT, "tests", "golden", "golden_geometry.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def golden_sims():
    import json

    with open(os.path.join(ROOT, "tests", "golden", "golden_sims.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def ref_inputs():
    return os.path.join(ROOT, "tests", "golden", "ref_inputs")
