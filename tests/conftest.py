"""Shared pytest configuration.

Markers:
    gpu  -- needs a HIP device (run with `-m gpu`); everything else runs on CPU.

The product library (prism-q_amd/libprism_md.so, libprism_gen.so) and the
oracle (oracle/liboracle.so) are built by __graft_entry__.build(); the tests
load what is built and fail loudly when a library is missing.
"""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "prism-q_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); run with -m gpu")


@pytest.fixture(scope="session")
def pm():
    import prism_md
    return prism_md


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    pyoracle.lib()  # fail loudly if liboracle.so is not built
    return pyoracle


@pytest.fixture(scope="session")
def small_top(pm):
    """192-atom generated chain (the oracle_small.npz system)."""
    return pm.Topology.generate(192, seed=11)


@pytest.fixture(scope="session")
def rng():
    return np.random.default_rng(12345)
