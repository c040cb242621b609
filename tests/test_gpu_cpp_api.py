"""The reference's own unit tests of the RD path (src/test/pdesolver.cpp:7-57, src/test/grad.cpp:7-80),
written in C++ against the reference's class interface in include/glia/ and run on the GPU, plus
that interface against the C ABI and central differences (tests/cpp/test_reference_api.cpp)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "test_reference_api.bin")


def test_reference_unit_tests_through_cpp_interface():
    assert os.path.exists(BIN), "build first: make -C glia_amd/csrc"
    print("case C++ interface: pdesolver KAT, grad objective, ABI and FD checks", flush=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=150)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout[-2000:]
    assert "0 failed" in r.stdout
