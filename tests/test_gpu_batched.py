"""Many uneven batches stepped by one batch plan (hepem_batch_plan_*):
bit-identical to hepem_step on each batch, before and after tuning the
launch shape, and when the run is captured into a hipGraph and replayed."""
import math

import numpy as np
import pytest

from conftest import same_bits, valid_secondaries
from oracle.oracle_py import OUT_FIELDS, TRACK_FIELDS
from test_gpu_parity import gpu_step

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

SEED = 0xBA7C4


def _print(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU test run without a visible GPU"
    return torch.device("cuda:0")


@pytest.fixture(scope="module")
def ctx64(tables2, dev):
    from g4hepem_amd import HEPEM_FP64, Context
    return Context(tables2, HEPEM_FP64)


def make_batches(sizes, dev, seed0=900):
    from g4hepem_amd.tracks import StepOutSoA, TrackSoA, step_tracks
    host, dev_batches = [], []
    first = 0
    for k, n in enumerate(sizes):
        tr = step_tracks(n, seed0 + k, 4, lnemin=math.log(1e-2), lnemax=math.log(1e5), first_id=first)
        first += n
        if n > 10:
            tr["imc"][3] = 77     # one rejected track per batch
        host.append(tr)
        dev_batches.append((TrackSoA.from_numpy(tr, "torch", dev), TrackSoA(n, "torch", dev), StepOutSoA(n, "torch", dev)))
    return host, dev_batches


def check_against_unbatched(ctx, host, dev_batches, dev, label):
    for k, (tr, (tin, tout, res)) in enumerate(zip(host, dev_batches)):
        ref_t, ref_r, ref_rej = gpu_step(ctx, SEED, tr, dev)
        got_t, got_r = tout.to_numpy(), valid_secondaries(res.to_numpy())
        for f in OUT_FIELDS:
            assert same_bits(got_r[f], ref_r[f]), (label, k, f)
        for f in TRACK_FIELDS:
            if f.startswith("msc"):
                m = np.tile(tr["ptype"] != 2, TRACK_FIELDS[f][1])
                assert same_bits(got_t[f][m], ref_t[f][m]), (label, k, f)
            else:
                assert same_bits(got_t[f], ref_t[f]), (label, k, f)


def test_batched_equals_unbatched(ctx64, dev):
    from g4hepem_amd import BatchPlan
    sizes = [1, 63, 64, 65, 1000, 4095, 4097, 30000, 0, 100000, 257]
    host, batches = make_batches(sizes, dev)
    plan = BatchPlan(ctx64, batches)
    _print("batched run, default shape", plan.shape())
    plan.run(SEED)
    torch.cuda.synchronize()
    for k, n in enumerate(sizes):
        assert plan.rejected(k) == (1 if n > 10 else 0), k
    check_against_unbatched(ctx64, host, batches, dev, "default")


def test_batched_tune_then_run(ctx64, dev):
    from g4hepem_amd import BatchPlan
    rng = np.random.Generator(np.random.Philox(5))
    sizes = np.exp(rng.uniform(math.log(1e2), math.log(3e4), 24)).astype(int).tolist()
    host, batches = make_batches(sizes, dev, seed0=950)
    plan = BatchPlan(ctx64, batches)
    block, grid = plan.tune(SEED)
    shape = plan.get_shape()
    _print("tuned shape", shape)
    assert block == shape["block"] and grid == shape["grid"]
    assert (block, shape["sort_tile"], shape["persistent"]) in BatchPlan.shapes()
    if not shape["persistent"] and shape["sort_tile"] and block == 256:
        # the tile-local particle partition: 16 blocks per 4096-slot tile,
        # tiles rounded up to a multiple of the 8 XCDs
        ntiles = sum((n + 4095) // 4096 for n in sizes)
        assert grid == (ntiles + 7) // 8 * 8 * 16
    elif not shape["persistent"]:
        assert grid == sum((n + block - 1) // block for n in sizes)
    for _, tout, res in batches:  # results must come from the run below
        res["edep"].fill_(-1.0)
    plan.run(SEED)
    torch.cuda.synchronize()
    check_against_unbatched(ctx64, host, batches, dev, "tuned")


def test_batched_every_shape_bit_identical(ctx64, dev):
    """Every launch shape the tuner may choose (block size, process sort on
    and off, each sort tile size, dispatched and persistent) gives the
    unbatched results bit for bit, on batches whose sizes straddle the tile
    sizes."""
    from g4hepem_amd import BatchPlan
    from g4hepem_amd.tracks import StepOutSoA, TrackSoA, step_outputs_differ
    sizes = [1, 1023, 1024, 1025, 2049, 8191, 8193, 20000, 0, 64, 4096, 300]
    host, batches = make_batches(sizes, dev, seed0=1200)
    work = ctx64.work_buffer(max(sizes), dev)
    refs = []
    for tin, _, _ in batches:
        t_ref, r_ref = TrackSoA(tin.n, "torch", dev), StepOutSoA(tin.n, "torch", dev)
        ctx64.step(SEED, tin, t_ref, r_ref, work)
        refs.append((t_ref, r_ref))
    plan = BatchPlan(ctx64, batches)
    shapes = BatchPlan.shapes()
    assert len(shapes) == 30
    for block, tile, pers in shapes:
        got = plan.set_shape(block, tile, pers)
        assert (got["block"], got["sort_tile"], got["persistent"]) == (block, tile, pers)
        for _, tout, res in batches:   # the checked results must come from this run
            res["winner"].fill_(-7)
            tout["ekin"].fill_(-1.0)
        plan.run(SEED)
        torch.cuda.synchronize()
        for k, ((tin, tout, res), (t_ref, r_ref)) in enumerate(zip(batches, refs)):
            diff = step_outputs_differ(tin["ptype"], tout, res, t_ref, r_ref)
            assert not diff, (block, tile, pers, k, diff)
        assert plan.rejected(7) == 1
    _print(f"{len(shapes)} shapes bit-identical")
    with pytest.raises(RuntimeError):
        plan.set_shape(96, 4096, 0)
    with pytest.raises(RuntimeError):
        plan.set_shape(256, 3000, 0)


@pytest.mark.parametrize("form", ["fp64", "fp32"])
def test_batched_tile_local_layouts(tables2, dev, form):
    """The sorted, dispatched 256-thread shape steps over the tile-local
    particle partition of the plan's tiles (numbered across the batches): on
    batches with shuffled, e-/e+-only and gamma-only layouts, sizes below,
    at and across the 4096-slot tile, and more tiles than one XCD share's
    tail, every batch equals hepem_step bit for bit."""
    from g4hepem_amd import HEPEM_FP32T, HEPEM_FP64, BatchPlan, Context
    from g4hepem_amd.tracks import StepOutSoA, TrackSoA, step_outputs_differ
    ctx = Context(tables2, HEPEM_FP64 if form == "fp64" else HEPEM_FP32T)
    sizes = [4096, 1, 70000, 4097, 0, 300000, 63, 12288, 9000]
    host, batches = make_batches(sizes, dev, seed0=1400)
    rng = np.random.default_rng(14)
    for k, (tr, (tin, _, _)) in enumerate(zip(host, batches)):
        if k % 3 == 0:
            tr["ptype"][:] = rng.permutation(tr["ptype"])
        elif k % 3 == 1:
            tr["ptype"][:] = np.where(rng.random(len(tr["ptype"])) < 0.7, 0, 1).astype(np.int8)
        else:
            tr["ptype"][:] = 2
        tin["ptype"].copy_(torch.from_numpy(tr["ptype"]))
    work = ctx.work_buffer(max(sizes), dev)
    refs = []
    for tin, _, _ in batches:
        t_ref, r_ref = TrackSoA(tin.n, "torch", dev), StepOutSoA(tin.n, "torch", dev)
        ctx.step(SEED, tin, t_ref, r_ref, work)
        refs.append((t_ref, r_ref))
    plan = BatchPlan(ctx, batches)
    for tile in (1024, 4096, 8192):
        _print(f"tile-local batched, {form}, sort tile {tile}")
        got = plan.set_shape(256, tile, 0)
        ntiles = sum((n + 4095) // 4096 for n in sizes)
        assert got["grid"] == (ntiles + 7) // 8 * 8 * 16
        for _, tout, res in batches:
            res["winner"].fill_(-7)
            tout["ekin"].fill_(-1.0)
        plan.run(SEED)
        torch.cuda.synchronize()
        for k, ((tin, tout, res), (t_ref, r_ref)) in enumerate(zip(batches, refs)):
            diff = step_outputs_differ(tin["ptype"], tout, res, t_ref, r_ref)
            assert not diff, (form, tile, k, diff)
    plan.close()
    ctx.close()


def test_batched_graph_replay(ctx64, dev):
    """The plan run captured into a graph; every replay, with eager work and
    synchronisation in between, reproduces the eager result."""
    from g4hepem_amd import BatchPlan
    sizes = [5000, 123, 70000, 9999, 1]
    host, batches = make_batches(sizes, dev, seed0=990)
    plan = BatchPlan(ctx64, batches)
    plan.run(SEED)
    torch.cuda.synchronize()
    eager = [(tout.to_numpy(), valid_secondaries(res.to_numpy())) for _, tout, res in batches]
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            plan.run(SEED, stream=s)
    torch.cuda.synchronize()
    for rep in range(3):
        _print(f"graph replay {rep}")
        for _, tout, res in batches:
            tout["ekin"].zero_()
            res["edep"].fill_(-1.0)
            res["winner"].fill_(-7)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for k, (_, tout, res) in enumerate(batches):
            got_t, got_r = tout.to_numpy(), valid_secondaries(res.to_numpy())
            for f in OUT_FIELDS:
                assert same_bits(got_r[f], eager[k][1][f]), (rep, k, f)
            assert same_bits(got_t["ekin"], eager[k][0]["ekin"]), (rep, k)
        assert plan.rejected(2) == 1


@pytest.mark.parametrize("sizes", [[1, 63, 64, 65, 1000, 4097, 30000, 0, 257],
                                   [70000, 5, 20000]],
                         ids=["ragged", "above-65536"])
def test_batched_equals_unbatched_fp32(tables2, dev, sizes):
    """The FP32-table form of a batch plan (the automatic sort choice takes the
    two-stage kernels) is bit-identical to hepem_step on each batch."""
    from g4hepem_amd import HEPEM_FP32T, BatchPlan, Context
    ctx32 = Context(tables2, HEPEM_FP32T)
    host, batches = make_batches(sizes, dev, seed0=990)
    plan = BatchPlan(ctx32, batches)
    plan.run(SEED)
    torch.cuda.synchronize()
    check_against_unbatched(ctx32, host, batches, dev, "fp32")
