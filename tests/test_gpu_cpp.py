"""The C++ host layers on the GPU: the in-tree test programs (built by
`python g4hepem_amd/build.py cpp`, part of build()):
  tests/cpp/test_managers (include/g4hepem_amd.hpp): ElectronManager /
    GammaManager / Stepper HowFar+Perform equal the fused step bit for bit,
    wrong-kind tracks are rejected, errors are exceptions carrying the C-ABI
    status;
  tests/cpp/test_onetrack (include/g4hepem_amd_onetrack.hpp): the one-track
    G4HepEmElectronManager / G4HepEmGammaManager HowFar, a geometry cut and
    Perform on a G4HepEmTLData equal the same tracks stepped as one batch, bit
    for bit."""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

CPP = Path(__file__).resolve().parent / "cpp"


@pytest.mark.parametrize("name", ["test_managers", "test_onetrack"])
def test_cpp_host_layer(name):
    binary = CPP / name
    assert binary.exists(), f"{binary} is not built (python g4hepem_amd/build.py cpp)"
    p = subprocess.run([str(binary)], capture_output=True, text=True, timeout=240)
    print(p.stdout, flush=True)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "ALL OK" in p.stdout
