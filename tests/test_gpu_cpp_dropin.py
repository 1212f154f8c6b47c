"""The reference's FunctionalityTests matrix, ported to C++ against the drop-in
header include/KnnTrees.hpp (tests/cpp/functionality_tests.cpp), run on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "functionality_tests")


def test_cpp_dropin_functionality_matrix():
    assert os.path.exists(BIN), "build the C++ test with python -m cudaexactknn_amd.build"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=110)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "All tests passed" in r.stdout
