"""The C++ drop-in surface (include/clique.hpp) driven like the reference's
tests/Solve.cpp, on the GPU through libcliqg.so."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "solve_driver")


@pytest.mark.gpu
@pytest.mark.parametrize("args", [["30", "30", "30", "5"],               # config 1 shape
                                  ["17", "9", "13", "3", "fp32"],
                                  ["12", "12", "12", "2", "fp64", "complex"],
                                  ["12", "12", "12", "2", "fp32", "complex"],
                                  ["16", "14", "12", "3", "fp64", "real", "block"],
                                  ["12", "12", "12", "2", "fp32", "complex", "block"],
                                  ["14", "13", "12", "2", "fp64", "complex", "ldl", "2d"],
                                  ["12", "11", "10", "2", "fp64", "real", "block", "2d"],
                                  ["13", "12", "11", "3", "fp64", "complex", "intrapiv"],
                                  ["14", "12", "10", "2", "fp32", "real", "intrapiv"],
                                  ["11", "12", "13", "2", "fp64", "real", "intrapiv", "2d"]])
def test_cpp_driver(device, args):
    if not os.path.exists(DRIVER):
        pytest.fail("tests/cpp/solve_driver is not built (run __graft_entry__.build())")
    r = subprocess.run([DRIVER] + args, capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and "PASSED" in r.stdout, r.stdout + r.stderr
