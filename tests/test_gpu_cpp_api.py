"""The C++ builder mirror (cpp/differential_equations.hpp) running the
reference's accuracy / interpolation / error tests on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_cpp_builder_runs_reference_tests():
    exe = os.path.join(HERE, "cpp", "_build", "test_accuracy")
    if not os.path.exists(exe):
        subprocess.check_call(["make", "-s", "-C", os.path.join(HERE, "cpp")])
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "ALL PASSED" in out.stdout
    assert "point outputs passed" in out.stdout
