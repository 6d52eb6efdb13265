"""Shared pytest setup: the `gpu` marker and cached host graphs."""

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libgapgpu.so)")
    config.addinivalue_line("markers", "slow: full-size workloads")


_GRAPHS = {}


def cached_graph(key, make):
    if key not in _GRAPHS:
        _GRAPHS[key] = make()
    return _GRAPHS[key]


@pytest.fixture(scope="session")
def golden():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "golden.json")) as f:
        return json.load(f)
