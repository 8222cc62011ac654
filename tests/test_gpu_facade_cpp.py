"""Runs the C++ facade test (tests/cpp/facade_test.cpp): FeCl's own LDPC test cases written
against include/fecl/Ldpc.hpp, linked to the HIP library."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "cpp", "facade_test")


@pytest.mark.gpu
def test_cpp_facade_reference_cases():
    assert os.path.exists(EXE), "build it with `make -C tests/cpp` (done by __graft_entry__.build())"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
