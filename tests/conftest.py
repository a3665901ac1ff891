import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def kernel_fx():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "ref_kernel_fixtures.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def matrix_fx():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "ref_matrix_fixtures.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def gpu():
    """Skips unless a GPU is visible; fails loudly if the HIP library is missing."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from edge_amd import _lib
    _lib.lib()  # raises when libedge_amd.so is missing
    return torch.device("cuda:0")


@pytest.fixture(scope="session")
def visco_fx():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "ref_visco_fixtures.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def rupture_fx():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "ref_rupture_fixtures.json")) as f:
        return json.load(f)
