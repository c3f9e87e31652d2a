"""Shared test configuration.

`gpu` tests need a HIP device and the built libbqp.so; everything else runs on
the CPU (the oracle is built on demand with gcc).
"""
import os
import sys

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libbqp.so")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle as O
    O.lib()
    return O
