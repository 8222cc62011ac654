"""Runs the C++ tests: FeCl's own LDPC test cases written against include/fecl/Ldpc.hpp
(tests/cpp/facade_test.cpp), and the MATLAB gateway driven through a test double of the MEX
API (tests/cpp/mex_gateway_test.cpp), both linked to the HIP library."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "cpp", "facade_test")


@pytest.mark.gpu
def test_cpp_facade_reference_cases():
    assert os.path.exists(EXE), "build it with `make -C tests/cpp`"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.gpu
def test_matlab_gateway_through_the_mex_test_double():
    """fec.bin.wrap's function ids and conventions over the HIP decoder: construct from a sparse
    DvbS2 matrix and from a dense Gallager matrix, sizes, encode/check/decode/soDecode against
    the C++ API, decoder options, work group size, puncturing, save/load, destroy."""
    exe = os.path.join(HERE, "cpp", "mex_gateway_test")
    assert os.path.exists(exe), "build it with `make -C tests/cpp`"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
