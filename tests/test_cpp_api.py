"""The C++ API header (include/lofloat/lo_float.hpp) over the C-ABI: tests/cpp/test_api.cpp,
built by __graft_entry__.build().  The CPU run checks argument handling only; the GPU run
checks round_fp_params / round_mantissa / run_virtual_mx_round and Gemm in all eight
A/B/D layout combinations (with and without beta * C) bit-exactly against the oracle."""
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(REPO, "tests", "cpp", "test_api")


def _exe():
    if not os.path.exists(EXE):
        import __graft_entry__
        __graft_entry__.build_cpp_tests()
    return EXE


def test_cpp_api_cpu():
    r = subprocess.run([_exe(), "cpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.gpu
def test_cpp_api_gpu():
    r = subprocess.run([EXE, "gpu"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
