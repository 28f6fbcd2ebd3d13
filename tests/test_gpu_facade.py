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
