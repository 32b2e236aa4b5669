import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built HIP library")


@pytest.fixture(scope="session")
def tiny():
    from mc_kernel_amd import synth
    return synth.make_problem("tiny")


@pytest.fixture(scope="session")
def tiny_oracle(tiny):
    from oracle import oracle
    return oracle.Oracle(tiny.setup)
