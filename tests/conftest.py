import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the device library)")


@pytest.fixture(scope="session")
def mats():
    from opus_destruct_2_amd import gen
    return gen.materials()


@pytest.fixture(scope="session")
def ctx():
    from opus_destruct_2_amd import stress
    return stress.default_context()
