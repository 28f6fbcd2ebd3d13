"""The C++ StressProblem over the C ABI (include/svxg_stress.hpp) passes the reference's own stress
tests (tests/core/test_stress.cpp) restated in tests/cpp/test_stress_facade.cpp."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_stress_facade")


def test_cpp_facade_restated_reference_tests():
    assert os.path.exists(EXE), "built by __graft_entry__.build()"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=240)
    print(r.stdout[-4000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "0 failed" in r.stdout


LOOPS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_world_loops")


def test_cpp_world_loops_on_the_block():
    """World's judge and probe loops (a bond_load and a bond_utilization a bond) compiled unchanged
    against the header with a world-shaped material table: phi bit-identical to the device's checks,
    one device pass a loop over the 64 x 64 x 48 block's 422,375 bonds."""
    assert os.path.exists(LOOPS), "built by __graft_entry__.build()"
    r = subprocess.run([LOOPS], capture_output=True, text=True, timeout=240)
    print(r.stdout[-4000:])
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert "0 failed" in r.stdout
