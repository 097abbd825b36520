"""HIP stepping on tables with lepto- and photo-nuclear cross sections
(Tables.set_nuclear), so that e-/e+ winner 3 and gamma process 3 occur,
against the golden vectors the compiled reference wrote on those tables and
against the oracle at the full config-2 shape, through the C-ABI.

Reference code on this path: G4HepEmElectronManager.icc:85,418-421,534-541,
570-574; G4HepEmGammaManager.icc:199-214."""
import math

import numpy as np
import pytest

from conftest import CUT, SET2, STEP_SEED, golden_tracks, same_bits
from oracle.oracle_py import OUT_FIELDS, TRACK_FIELDS
from test_gpu_parity import compare_step, gpu_step

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

GOLDEN_NUC = __import__("pathlib").Path(__file__).resolve().parent / "golden" / "golden_nuc_v1.npz"
NUC_SCALES = (0.5, 0.2)


def _print(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU test run without a visible GPU"
    return torch.device("cuda:0")


@pytest.fixture(scope="module")
def gnuc():
    return np.load(GOLDEN_NUC, allow_pickle=False)


@pytest.fixture(scope="module")
def tnuc():
    from g4hepem_amd import Tables
    return Tables(SET2, CUT).set_nuclear(*NUC_SCALES)


@pytest.fixture(scope="module")
def cnuc(tnuc, dev):
    from g4hepem_amd import HEPEM_FP64, Context
    return Context(tnuc, HEPEM_FP64)


@pytest.mark.parametrize("r", [0, 1])
def test_step_nuclear_golden(gnuc, cnuc, dev, r):
    _print(f"nuclear step golden regime {r}")
    tin = golden_tracks(gnuc, f"st{r}_in")
    t, res, nrej = gpu_step(cnuc, STEP_SEED + 4, tin, dev)
    assert nrej == 0
    ref_t = {k: gnuc[f"st{r}_out_{k}"] for k in TRACK_FIELDS}
    ref_r = {k: gnuc[f"st{r}_res_{k}"] for k in OUT_FIELDS}
    compare_step(tin, t, res, ref_t, ref_r, label=f"nuc regime{r}")
    el = tin["ptype"] != 2
    assert ((res["winner"] == 3) & el).sum() >= 0.01 * el.sum()
    assert ((res["winner"] == 3) & ~el).sum() >= 0.01 * (~el).sum()


def test_step_chain_nuclear_golden(gnuc, cnuc, dev):
    cur = golden_tracks(gnuc, "ch_in")
    for s in range(3):
        _print(f"nuclear step chain {s}")
        alive = cur["ekin"] > 0
        nxt, res, nrej = gpu_step(cnuc, STEP_SEED + 5, cur, dev)
        assert nrej == (~alive).sum()
        ref_r = {k: gnuc[f"ch{s}_res_{k}"] for k in OUT_FIELDS}
        ref_t = dict(nxt)
        for k in ("ekin", "dir", "num_ia_left", "rng_ctr", "rng_gauss", "rng_has_gauss", "msc_init_range",
                  "msc_dyn_range_factor", "msc_tlimit_min", "msc_first_step"):
            ref_t[k] = gnuc[f"ch{s}_out_{k}"]
        compare_step(cur, nxt, res, ref_t, ref_r, alive=alive, label=f"nuc chain{s}")
        cur = dict(nxt)
        for k in ref_t:
            cur[k] = ref_t[k]


@pytest.mark.parametrize("sort", [True, False])
def test_step_nuclear_parity_1M(tnuc, oracle, dev, sort):
    """Config-2 shape on the nuclear tables, with and without the process sort."""
    from g4hepem_amd import HEPEM_FP64, Context
    from g4hepem_amd.tracks import step_tracks
    ctx = Context(tnuc, HEPEM_FP64)
    ctx.set_process_sort(sort)
    n = 1_000_000
    tr = step_tracks(n, 20261, 4)
    _print(f"nuclear step parity 1M sort={sort}: gpu")
    t, res, nrej = gpu_step(ctx, 0xBEEF, tr, dev)
    assert nrej == 0
    _print("  oracle")
    ref_t, ref_r = oracle.step(tnuc, 0xBEEF, tr, nthreads=16)
    compare_step(tr, t, res, ref_t, ref_r, label="nuc 1M")
    el = tr["ptype"] != 2
    w = res["winner"]
    _print(f"  winner 3: e-/e+ {((w == 3) & el).mean() / el.mean():.4f}, gamma {((w == 3) & ~el).mean() / (~el).mean():.4f}")
    assert ((w == 3) & el).sum() >= 0.01 * el.sum()
    assert ((w == 3) & ~el).sum() >= 0.01 * (~el).sum()


def test_split_equals_fused_nuclear(tnuc, cnuc, dev):
    """HowFar + Perform as separate calls equal the fused step on the nuclear tables."""
    from g4hepem_amd.tracks import HowFarSoA, StepOutSoA, TrackSoA, step_tracks
    tr = step_tracks(200000, 31, 4, lnemin=math.log(1.0), lnemax=math.log(1e6))
    a_t, a_r, _ = gpu_step(cnuc, 5, tr, dev)
    t = TrackSoA.from_numpy(tr, "torch", dev)
    hs = HowFarSoA(len(tr["ekin"]), "torch", dev)
    res = StepOutSoA(len(tr["ekin"]), "torch", dev)
    work = cnuc.work_buffer(len(tr["ekin"]), dev)
    cnuc.howfar(5, t, hs, work)
    cnuc.perform(5, t, hs, res, work)
    torch.cuda.synchronize()
    from conftest import valid_secondaries
    b_r = valid_secondaries(res.to_numpy())
    b_t = t.to_numpy()
    for k in OUT_FIELDS:
        assert same_bits(a_r[k], b_r[k]), k
    for k in ("ekin", "dir", "num_ia_left", "rng_ctr", "rng_gauss"):
        assert same_bits(a_t[k], b_t[k]), k
    assert (a_r["winner"] == 3).sum() > 1000


def test_fp32_nuclear_decisions(tnuc, oracle, dev):
    """FP32-table form on the nuclear tables: the nuclear winners follow the
    FP64 oracle like the other decisions."""
    from g4hepem_amd import HEPEM_FP32T, Context
    from g4hepem_amd.tracks import step_tracks
    ctx = Context(tnuc, HEPEM_FP32T)
    tr = step_tracks(200000, 45, 4, lnemin=math.log(1.0), lnemax=math.log(1e5))
    t, res, nrej = gpu_step(ctx, 6, tr, dev)
    assert nrej == 0
    ref_t, ref_r = oracle.step(tnuc, 6, tr, nthreads=16)
    agree = (res["winner"] == ref_r["winner"]).mean()
    n3 = (ref_r["winner"] == 3).sum()
    both3 = ((res["winner"] == 3) & (ref_r["winner"] == 3)).sum()
    _print(f"  fp32 winner agreement {agree:.5f}; winner-3 {both3}/{n3}")
    assert agree > 0.99 and n3 > 2000 and both3 >= 0.99 * n3
