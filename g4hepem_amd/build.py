"""Build the native libraries in-tree.

* ``g4hepem_amd/libg4hepem_amd.so`` -- the product: HIP kernels for gfx950,
  the C-ABI (include/g4hepem_amd.h) and the offline table builder
  (include/g4hepem_amd_tables.h), compiled with hipcc.
* ``oracle/liboracle.so`` -- the C restatement used as test oracle, and
  ``oracle/_ref/libhepem_ref.so`` -- the reference's own sources compiled
  from the reference tree (only where that tree exists), both via
  ``oracle/Makefile``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "g4hepem_amd"
CSRC = PKG / "csrc"
LIB = PKG / "libg4hepem_amd.so"
REFERENCE_TREE = Path("/root/reference")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # round every product and sum as written: the FP64 form must reproduce
    # the host reference's arithmetic
    "-ffp-contract=off",
    # no packed FP32 instructions (v_pk_*_f32) in generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, cwd=None):
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], cwd=cwd, check=True)


def build_native() -> Path:
    sources = [CSRC / "hepem_capi.hip", CSRC / "hepem_shower.hip", CSRC / "hepem_stack.hip", CSRC / "tables" / "hepem_tables.cpp",
               CSRC / "tables" / "hepem_json.cpp", CSRC / "tables" / "hepem_validate.cpp"]
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [_hipcc(), *HIPCC_FLAGS, f"-I{ROOT / 'include'}", f"-I{CSRC}", "-o", tmp, *sources]
    _run(cmd)
    os.replace(tmp, LIB)
    return LIB


CPP_TEST_DIR = ROOT / "tests" / "cpp"
# the C++ host-layer test programs: g4hepem_amd.hpp (batches) and
# g4hepem_amd_onetrack.hpp (the one-track drop-in)
CPP_TESTS = ("test_managers", "test_onetrack")
CPP_TEST_BIN = CPP_TEST_DIR / "test_managers"


def build_cpp_tests() -> Path:
    """The C++ host-layer test programs, linked against the in-tree library."""
    for name in CPP_TESTS:
        out = CPP_TEST_DIR / name
        tmp = out.with_suffix(".tmp")
        cmd = [_hipcc(), "--offload-arch=gfx950", "-O2", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
               "-packed-fp32-ops", f"-I{ROOT / 'include'}", "-o", tmp, CPP_TEST_DIR / f"{name}.cpp", f"-L{PKG}",
               "-lg4hepem_amd", "-Wl,-rpath,$ORIGIN/../../g4hepem_amd"]
        _run(cmd)
        os.replace(tmp, out)
    return CPP_TEST_BIN


def build_oracle() -> None:
    make = shutil.which("make") or "make"
    _run([make, "-s", "-C", ROOT / "oracle", "all"])
    if REFERENCE_TREE.is_dir():
        _run([make, "-s", "-j8", "-C", ROOT / "oracle", "ref", f"REF={REFERENCE_TREE}"])


def build_all() -> None:
    build_native()
    build_cpp_tests()
    build_oracle()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    {"all": build_all, "native": build_native, "oracle": build_oracle, "cpp": build_cpp_tests}[what]()
