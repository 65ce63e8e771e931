import gzip
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI library)")


def load_ref_vectors(name):
    """Golden vectors from the reference's test/unit/*.json (copied as data)."""
    with gzip.open(os.path.join(GOLDEN, "ref", name + ".gz"), "rt") as f:
        return json.load(f)


@pytest.fixture(scope="session")
def ref_vectors():
    return load_ref_vectors
