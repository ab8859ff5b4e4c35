"""Test configuration: the `gpu` marker, import paths, and shared fixtures.

`-m "not gpu"` runs the oracle checks, host logic and ABI-loading tests on the
CPU; `-m gpu` runs the HIP library against the oracle on an MI355X."""
import os
import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
for sub in ("oracle", "shakti.jl_amd"):
    p = str(ROOT / sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built libshakti_hip.so")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def hip():
    import shakti_hip
    shakti_hip.load()
    return shakti_hip
