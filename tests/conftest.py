import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")
INPUTS = os.path.join(GOLDEN_DIR, "inputs")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (runs on the GPU box)")


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(GOLDEN_DIR, "golden.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def fixture_graphs():
    import oracle as O
    return {name: O.read_adjacency_graph(os.path.join(INPUTS, name))
            for name in ("rMatGraph_J_5_100", "triangles.txt", "star.txt")}
