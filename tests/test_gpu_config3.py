"""The 64-batch workload at its real shape (sizes log-uniform in [1e3, 1e6],
about 6.45M tracks, built exactly as bench.py builds it): one tuned batch plan
in each form is bit-identical to hepem_step on every batch, and in FP64 a
20k-track slice of four batches matches the oracle."""
import numpy as np
import pytest

from conftest import STEP_SEED, valid_secondaries
from test_gpu_parity import compare_step

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _print(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU test run without a visible GPU"
    return torch.device("cuda:0")


@pytest.fixture(scope="module")
def workload(dev):
    import bench
    from g4hepem_amd.tracks import StepOutSoA, TrackSoA
    host = bench.config3_batches(STEP_SEED)
    batches = [(TrackSoA.from_numpy(b, "torch", dev), TrackSoA(len(b["ekin"]), "torch", dev),
                StepOutSoA(len(b["ekin"]), "torch", dev)) for b in host]
    _print(f"config-3 workload: {len(host)} batches, {sum(len(b['ekin']) for b in host)} tracks")
    return host, batches


@pytest.mark.parametrize("precision", ["fp64", "fp32t"])
def test_config3_plan_bit_identical_to_unbatched(tables2, dev, workload, precision):
    import bench
    from g4hepem_amd import HEPEM_FP32T, HEPEM_FP64, BatchPlan, Context
    host, batches = workload
    assert len(batches) == 64 and sum(len(b["ekin"]) for b in host) > 6_000_000
    ctx = Context(tables2, HEPEM_FP64 if precision == "fp64" else HEPEM_FP32T)
    plan = BatchPlan(ctx, batches)
    shape = plan.tune(STEP_SEED)
    _print(f"{precision}: tuned shape {shape}")
    for _, _, res in batches:   # the checked results must come from the run below
        res["winner"].fill_(-7)
    plan.run(STEP_SEED)
    torch.cuda.synchronize()
    chk = bench.verify_batches(ctx, batches, STEP_SEED)
    assert chk["bit_identical_to_unbatched"], chk["mismatches"]
    assert chk["batches_checked"] == 64
    if precision == "fp64":
        from oracle.oracle_py import Oracle
        orc = Oracle()
        for k in (0, 21, 42, 63):
            tr, (_, tout, res) = host[k], batches[k]
            m = min(5000, len(tr["ekin"]))
            n = len(tr["ekin"])
            sl = {f: np.concatenate([v[c * n: c * n + m] for c in range(len(v) // n)]) for f, v in tr.items()}
            g_t, g_r = tout.to_numpy(), res.to_numpy()
            g_t = {f: np.concatenate([v[c * n: c * n + m] for c in range(len(v) // n)]) for f, v in g_t.items()}
            g_r = {f: np.concatenate([v[c * n: c * n + m] for c in range(len(v) // n)]) for f, v in g_r.items()}
            ref_t, ref_r = orc.step(tables2, STEP_SEED, sl, nthreads=8)
            frac = compare_step(sl, g_t, valid_secondaries(g_r), ref_t, ref_r, label=f"batch {k}")
            _print(f"batch {k}: {m} tracks against the oracle, bit-exact float fraction {frac:.4f}")
    plan.close()
    ctx.close()
