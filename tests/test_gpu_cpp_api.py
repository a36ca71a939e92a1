"""Runs the C++ API test program (tests/cpp/api_test.cpp): the gs_native classes used the way
the reference's integration tests use them (initialise, train, density boundary, save/reload,
class-level density, solver and KNN cases)."""
import pathlib
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BINARY = pathlib.Path(__file__).resolve().parent / "cpp" / "api_test"


def test_cpp_api_program():
    assert BINARY.exists(), "tests/cpp/api_test is not built: run __graft_entry__.build()"
    result = subprocess.run([str(BINARY)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
    print(result.stdout)
    assert result.returncode == 0, result.stdout[-4000:]
