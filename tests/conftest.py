import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden" / "golden_v1.npz"
GOLDEN_GEOM = ROOT / "tests" / "golden" / "golden_geom_v1.npz"
SET1 = ["G4_Pb", "G4_lAr"]
SET2 = ["G4_PbWO4", "G4_Pb", "G4_lAr", "G4_Si"]
CUT = 0.7
STEP_SEED = 0x5EED2026


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def golden():
    return np.load(GOLDEN, allow_pickle=False)


@pytest.fixture(scope="session")
def golden_geom():
    return np.load(GOLDEN_GEOM, allow_pickle=False)


@pytest.fixture(scope="session")
def tables1():
    from g4hepem_amd import Tables
    return Tables(SET1, CUT)


@pytest.fixture(scope="session")
def tables2():
    from g4hepem_amd import Tables
    return Tables(SET2, CUT)


@pytest.fixture(scope="session")
def oracle():
    from oracle.oracle_py import Oracle
    return Oracle()


def golden_tracks(g, prefix):
    """Track state stored in a golden file; fields added after it was written
    (rng_stream: 0) get their neutral value."""
    from oracle.oracle_py import TRACK_FIELDS
    n = len(g[f"{prefix}_ekin"])
    out = {}
    for k, (dt, c) in TRACK_FIELDS.items():
        key = f"{prefix}_{k}"
        out[k] = g[key] if key in g.files else np.zeros(c * n, dtype=dt)
    return out


def valid_secondaries(res: dict) -> dict:
    """Results with the secondary slots at or beyond nsec cleared: those slots
    are not part of the result (the kernels leave them unwritten)."""
    out = dict(res)
    nsec = np.asarray(res["nsec"])
    n = len(nsec)
    if n == 0:
        return out
    for k, fill in (("sec_ptype", -1), ("sec_ekin", 0.0), ("sec_dir", 0.0)):
        a = np.array(res[k], copy=True)
        comps = len(a) // (2 * n)
        v = a.reshape(comps, 2, n)
        for s in range(2):
            v[:, s, nsec <= s] = fill
        out[k] = v.reshape(-1)
    return out


def same_bits(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.dtype.kind == "f":
        return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))
    return np.array_equal(a, b)
