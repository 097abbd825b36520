"""The C++ host layer (include/g4hepem_amd.hpp) on the GPU: runs the
in-tree test program tests/cpp/test_managers (built by
`python g4hepem_amd/build.py cpp`, part of build()), which checks that
ElectronManager / GammaManager / Stepper HowFar+Perform equal the fused step
bit for bit, that wrong-kind tracks are rejected and that errors are
exceptions carrying the C-ABI status."""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

BIN = Path(__file__).resolve().parent / "cpp" / "test_managers"


def test_cpp_managers():
    assert BIN.exists(), f"{BIN} is not built (python g4hepem_amd/build.py cpp)"
    p = subprocess.run([str(BIN)], capture_output=True, text=True, timeout=240)
    print(p.stdout, flush=True)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "ALL OK" in p.stdout
