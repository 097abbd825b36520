"""HowFar and Perform as separate C-ABI calls with a caller geometry in
between (hepem_howfar / hepem_perform and the per-kind electron and gamma
entry points), against the fused step, the oracle and the golden vectors
the compiled reference wrote for geometry-limited steps.

Bar as for the fused step: integers exact, floats within 1e-12 relative in
the FP64 form; split == fused bit for bit in both forms."""
import math

import numpy as np
import pytest

from conftest import STEP_SEED, golden_tracks, same_bits, valid_secondaries
from oracle.oracle_py import OUT_FIELDS, TRACK_FIELDS
from test_gpu_parity import compare_step, gpu_step

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


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


@pytest.fixture(scope="module")
def ctx32(tables2, dev):
    from g4hepem_amd import HEPEM_FP32T, Context
    return Context(tables2, HEPEM_FP32T)


def gpu_split_step(ctx, seed, tracks: dict, dev, geom_limit=None, kind="mixed"):
    """howfar -> (geometry on the device: cut the geometrical step at
    geom_limit, set the post-step boundary flag) -> perform, in place."""
    from g4hepem_amd.tracks import HowFarSoA, StepOutSoA, TrackSoA
    n = len(tracks["ekin"])
    t = TrackSoA.from_numpy(tracks, "torch", dev)
    hs = HowFarSoA(n, "torch", dev)
    res = StepOutSoA(n, "torch", dev)
    work = ctx.work_buffer(n, dev)
    ctx.howfar(seed, t, hs, work, kind=kind)
    proposed = hs["geom_step"].clone()
    if geom_limit is not None:
        lim = torch.from_numpy(np.ascontiguousarray(geom_limit, np.float64)).to(dev)
        cut = lim < hs["geom_step"]
        hs["geom_step"].copy_(torch.where(cut, lim, hs["geom_step"]))
        t["on_boundary"].copy_(cut.to(torch.uint8))
    ctx.perform(seed, t, hs, res, work, kind=kind)
    torch.cuda.synchronize()
    return t.to_numpy(), valid_secondaries(res.to_numpy()), ctx.rejected(work), proposed.cpu().numpy()


@pytest.mark.parametrize("form", ["fp64", "fp32"])
def test_split_equals_fused(ctx64, ctx32, dev, form):
    from g4hepem_amd.tracks import step_tracks
    ctx = ctx64 if form == "fp64" else ctx32
    tr = step_tracks(200000, 61, 4, lnemin=math.log(1e-3), lnemax=math.log(1e6))
    tr["on_boundary"][:] = (np.arange(200000) % 5 == 2).astype(np.uint8)
    _print(f"split vs fused {form}")
    a_t, a_r, a_rej = gpu_step(ctx, 31, tr, dev, inplace=True)
    b_t, b_r, b_rej, _ = gpu_split_step(ctx, 31, tr, dev)
    assert a_rej == b_rej == 0
    for k in OUT_FIELDS:
        assert same_bits(a_r[k], b_r[k]), k
    el = np.tile(tr["ptype"] != 2, 1)
    for k, (dt, c) in TRACK_FIELDS.items():
        m = np.tile(el, c) if k.startswith("msc") else np.ones(c * len(el), bool)
        assert same_bits(a_t[k][m], b_t[k][m]), k


@pytest.mark.parametrize("sort", ["auto", "on"])
def test_split_geometry_golden(golden_geom, tables2, ctx64, dev, sort):
    """Geometry-limited 3-step chain against the compiled reference's vectors,
    with the default Perform (one kernel at this size) and with the process
    sort forced on."""
    from g4hepem_amd import HEPEM_FP64, Context
    assert tables2.digest() == str(golden_geom["digest_set2"])
    if sort == "on":
        ctx64 = Context(tables2, HEPEM_FP64)
        ctx64.set_process_sort(True)
    cur = golden_tracks(golden_geom, "gc_in")
    for s in range(3):
        _print(f"geometry chain step {s} (sort {sort})")
        alive = cur["ekin"] > 0
        lim = golden_geom[f"gc{s}_limit"]
        t, res, nrej, _ = gpu_split_step(ctx64, STEP_SEED + 2, cur, dev, geom_limit=lim)
        assert nrej == (~alive).sum()
        ref_r = {k: golden_geom[f"gc{s}_res_{k}"] for k in OUT_FIELDS}
        ref_t = dict(t)
        for k in ("ekin", "dir", "num_ia_left", "rng_ctr", "rng_gauss", "rng_has_gauss", "msc_init_range",
                  "msc_dyn_range_factor", "msc_tlimit_min", "msc_first_step", "on_boundary"):
            ref_t[k] = golden_geom[f"gc{s}_out_{k}"]
        compare_step(cur, t, res, ref_t, ref_r, alive=alive, label=f"geom{s}")
        cur = dict(t)
        for k in ref_t:
            cur[k] = ref_t[k]


def test_split_geometry_parity(tables2, ctx64, oracle, dev):
    from g4hepem_amd.tracks import step_tracks
    n = 300000
    tr = step_tracks(n, 62, 4, lnemin=math.log(1e-3), lnemax=math.log(1e7))
    lim = np.exp(np.random.Generator(np.random.Philox(63)).uniform(math.log(1e-4), math.log(20.0), n))
    _print("geometry parity 300k: gpu")
    t, res, nrej, proposed = gpu_split_step(ctx64, 32, tr, dev, geom_limit=lim)
    assert nrej == 0
    _print("geometry parity 300k: oracle")
    ref_t, ref_r = oracle.step(tables2, 32, tr, nthreads=16, geom_limit=lim)
    compare_step(tr, t, res, ref_t, ref_r, label="geometry")
    cut = lim < proposed
    assert 0.1 < cut.mean() < 0.9
    assert np.array_equal(t["on_boundary"], cut.astype(np.uint8))


def test_split_per_kind_entry_points(ctx64, dev):
    """electron-only and gamma-only batches through the per-kind calls equal
    the mixed calls; a slot of the other kind is rejected."""
    from g4hepem_amd.tracks import step_tracks
    tr = step_tracks(60000, 64, 4, lnemin=math.log(1e-2), lnemax=math.log(1e5))
    lim = np.exp(np.random.Generator(np.random.Philox(65)).uniform(math.log(1e-3), math.log(10.0), 60000))
    for kind, sel in (("electron", tr["ptype"] != 2), ("gamma", tr["ptype"] == 2)):
        sub = {k: v.reshape(TRACK_FIELDS[k][1], -1)[:, sel].ravel() for k, v in tr.items()}
        sl = lim[sel]
        _print(f"per-kind {kind}: {sel.sum()} tracks")
        a_t, a_r, a_rej, _ = gpu_split_step(ctx64, 33, sub, dev, geom_limit=sl, kind="mixed")
        b_t, b_r, b_rej, _ = gpu_split_step(ctx64, 33, sub, dev, geom_limit=sl, kind=kind)
        assert a_rej == b_rej == 0
        for k in OUT_FIELDS:
            assert same_bits(a_r[k], b_r[k]), (kind, k)
        for k in ("ekin", "dir", "num_ia_left", "rng_ctr", "rng_gauss", "on_boundary"):
            assert same_bits(a_t[k], b_t[k]), (kind, k)
    # a gamma in an electron-only batch and an electron in a gamma-only batch
    mixed = {k: v.reshape(TRACK_FIELDS[k][1], -1)[:, :20].ravel() for k, v in tr.items()}
    for kind, wrong in (("electron", mixed["ptype"] == 2), ("gamma", mixed["ptype"] != 2)):
        t, res, nrej, _ = gpu_split_step(ctx64, 34, mixed, dev, kind=kind)
        assert nrej == wrong.sum()
        assert np.all(res["winner"][wrong] == -99) and np.all(res["winner"][~wrong] != -99)
        assert same_bits(t["ekin"][wrong], mixed["ekin"][wrong])


@pytest.mark.parametrize("kind", ["mixed", "electron", "gamma"])
def test_split_process_sort_equals_unsorted(tables2, dev, kind):
    """Perform with the pending interactions sorted by process (three kernels)
    equals the single Perform kernel bit for bit, with a geometry limit and a
    rejected track."""
    from g4hepem_amd import HEPEM_FP64, Context
    from g4hepem_amd.tracks import step_tracks
    ctx = Context(tables2, HEPEM_FP64)
    n = 150000
    tr = step_tracks(n, 66, 4, lnemin=math.log(1e-3), lnemax=math.log(1e7))
    lim = np.exp(np.random.Generator(np.random.Philox(67)).uniform(math.log(1e-4), math.log(20.0), n))
    if kind != "mixed":
        sel = (tr["ptype"] != 2) if kind == "electron" else (tr["ptype"] == 2)
        tr = {k: v.reshape(TRACK_FIELDS[k][1], -1)[:, sel].ravel() for k, v in tr.items()}
        lim = lim[sel]
    tr["imc"][11] = 9  # one rejected track
    _print(f"split process sort {kind}: {len(tr['ekin'])} tracks")
    ctx.set_process_sort(True)
    a_t, a_r, a_rej, _ = gpu_split_step(ctx, 35, tr, dev, geom_limit=lim, kind=kind)
    ctx.set_process_sort(False)
    b_t, b_r, b_rej, _ = gpu_split_step(ctx, 35, tr, dev, geom_limit=lim, kind=kind)
    assert a_rej == b_rej == 1
    for k in OUT_FIELDS:
        assert same_bits(a_r[k], b_r[k]), (kind, k)
    for k in TRACK_FIELDS:
        assert same_bits(a_t[k], b_t[k]), (kind, k)
