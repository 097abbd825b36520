"""HIP kernels against the oracle and the golden vectors, through the C-ABI.

Bar: FP64 form bit-exact on every integer output and on the lookups; energies,
lengths and directions within 1e-12 relative (absolute for unit-vector
components) -- the only operations that may round differently from the host
are the device's sin, cos and pow, which need not round like glibc's.
"""
import math

import numpy as np
import pytest

from conftest import STEP_SEED, golden_tracks, same_bits, valid_secondaries
from oracle.oracle_py import OUT_FIELDS, TRACK_FIELDS

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

REL_TOL = 1e-12          # FP64 form: relative tolerance on floating outputs
ABS_TOL_DIR = 1e-12      # FP64 form: absolute tolerance on unit-vector components
INT_FIELDS = ("winner", "stopped", "nsec", "sec_ptype")


def _print(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU test run without a visible GPU"
    return torch.device("cuda:0")


@pytest.fixture(scope="module")
def ctx1(tables1, dev):
    from g4hepem_amd import HEPEM_FP64, Context
    return Context(tables1, HEPEM_FP64)


@pytest.fixture(scope="module")
def ctx2(tables2, dev):
    from g4hepem_amd import HEPEM_FP64, Context
    return Context(tables2, HEPEM_FP64)


@pytest.fixture(scope="module")
def ctx2_32(tables2, dev):
    from g4hepem_amd import HEPEM_FP32T, Context
    return Context(tables2, HEPEM_FP32T)


def gpu_lookup(ctx, pt, imc, e, le, dev):
    out = ctx.lookup(pt, torch.from_numpy(np.ascontiguousarray(imc, np.int32)).to(dev),
                     torch.from_numpy(np.ascontiguousarray(e)).to(dev), torch.from_numpy(np.ascontiguousarray(le)).to(dev))
    torch.cuda.synchronize()
    return out.cpu().numpy()


def gpu_step(ctx, seed, tracks: dict, dev, inplace=False):
    from g4hepem_amd.tracks import StepOutSoA, TrackSoA
    n = len(tracks["ekin"])
    tin = TrackSoA.from_numpy(tracks, "torch", dev)
    tout = tin if inplace else TrackSoA(n, "torch", dev)
    res = StepOutSoA(n, "torch", dev)
    work = ctx.work_buffer(n, dev)
    ctx.step(seed, tin, tout, res, work)
    torch.cuda.synchronize()
    return tout.to_numpy(), valid_secondaries(res.to_numpy()), ctx.rejected(work)


def compare_step(tin, t_gpu, r_gpu, t_ref, r_ref, alive=None, label=""):
    """Integers exact; floats within REL_TOL. Returns the bit-exact fraction of float outputs."""
    n = len(tin["ekin"])
    alive = np.ones(n, bool) if alive is None else alive
    el = (tin["ptype"] != 2) & alive
    r_gpu, r_ref = valid_secondaries(r_gpu), valid_secondaries(r_ref)
    exact, total = 0, 0
    for k, (dt, c) in OUT_FIELDS.items():
        m = np.tile(alive, c)
        a, b = r_gpu[k][m], r_ref[k][m]
        if k in INT_FIELDS:
            bad = np.nonzero(a != b)[0]
            assert len(bad) == 0, f"{label} {k}: {len(bad)} mismatches, first slots {bad[:5] % n}"
            continue
        exact += int((a.view(np.uint64) == b.view(np.uint64)).sum())
        total += a.size
        tol = ABS_TOL_DIR if k in ("sec_dir", "displacement") else 0.0
        err = np.abs(a - b) - tol
        lim = REL_TOL * np.abs(b)
        bad = np.nonzero(err > lim)[0]
        assert len(bad) == 0, f"{label} {k}: {len(bad)} beyond tolerance, e.g. {a[bad[:3]]} vs {b[bad[:3]]}"
    for k, (dt, c) in TRACK_FIELDS.items():
        sel = el if k.startswith("msc") else alive
        m = np.tile(sel, c)
        a, b = t_gpu[k][m], t_ref[k][m]
        if np.dtype(dt).kind != "f":
            assert np.array_equal(a, b), f"{label} state {k}"
            continue
        exact += int((a.view(np.uint64) == b.view(np.uint64)).sum())
        total += a.size
        tol = ABS_TOL_DIR if k == "dir" else 0.0
        bad = np.nonzero(np.abs(a - b) - tol > REL_TOL * np.abs(b))[0]
        assert len(bad) == 0, f"{label} state {k}: {len(bad)} beyond tolerance"
    return exact / max(total, 1)


# ---- config-1 lookups ------------------------------------------------------------------------------

@pytest.mark.parametrize("pt,name", [(0, "e"), (1, "p")])
def test_lookup_golden_bitexact(golden, ctx1, dev, pt, name):
    _print(f"lookup golden {name}")
    out = gpu_lookup(ctx1, pt, golden[f"c1_{name}_imc"], golden[f"c1_{name}_ekin"], golden[f"c1_{name}_lekin"], dev)
    assert same_bits(out, golden[f"c1_{name}_out"])


def test_lookup_parity_100k(tables1, ctx1, oracle, dev):
    from g4hepem_amd.tracks import lookup_queries
    imc, e, le = lookup_queries(100000, 2026, 2, margin=0.05)
    for pt in (0, 1):
        _print(f"lookup parity 100k ptype={pt}")
        ref = oracle.lookup(tables1, pt, imc, e, le)
        out = gpu_lookup(ctx1, pt, imc, e, le, dev)
        assert same_bits(out, ref)


@pytest.mark.parametrize("materials,form", [("set1", "fp64"), ("set1", "fp32"), ("eight", "fp64"), ("eight", "fp32")])
def test_lookup_staging_equals_global(tables1, oracle, dev, materials, form):
    """The lookups from tables staged in LDS equal those from the table image
    bit for bit, with the grid-stride loop over more queries than one pass of
    the staged grid; eight couples stage 56 kB in FP64, while FP32 needs
    84 kB, more than the staging limit, so it reads the image either way. FP64
    also matches the oracle bit for bit."""
    from g4hepem_amd import HEPEM_FP32T, HEPEM_FP64, Context, Tables
    from g4hepem_amd.tracks import lookup_queries
    tables = tables1 if materials == "set1" else Tables(["G4_Pb", "G4_lAr", "G4_Si", "G4_PbWO4", "G4_W", "G4_Cu", "G4_Fe",
                                                          "G4_WATER"], 0.7)
    nmc = 2 if materials == "set1" else 8
    ctx = Context(tables, HEPEM_FP64 if form == "fp64" else HEPEM_FP32T)
    imc, e, le = lookup_queries(700_001, 77, nmc, margin=0.05)
    imc[::1000] = -1  # invalid couples give NaN either way
    for pt in (0, 1):
        _print(f"lookup staging {materials} {form} ptype={pt}")
        ctx.set_lookup_staging(True)
        a = gpu_lookup(ctx, pt, imc, e, le, dev)
        ctx.set_lookup_staging(False)
        b = gpu_lookup(ctx, pt, imc, e, le, dev)
        assert same_bits(a, b)
        assert np.isnan(a[:, imc < 0]).all()
        if form == "fp64":
            ok = imc >= 0
            ref = oracle.lookup(tables, pt, imc[ok], e[ok], le[ok])
            assert same_bits(np.ascontiguousarray(a[:, ok]), ref)


def test_lookup_at_grid_nodes(tables1, ctx1, oracle, dev):
    """Energies on the loss-grid nodes and one ulp either side, where the
    range lands on (or next to) a bin edge of the inverse-range search: the
    hinted search returns the binary search's bin, so the lookups stay
    bit-exact against the oracle."""
    from g4hepem_amd.data_layout import arr, view
    ed = view(tables1).fTheElectronData.contents
    grid = arr(ed.fELossEnergyGrid, ed.fELossEnergyGridSize)
    e = np.concatenate([grid, np.nextafter(grid, 0.0), np.nextafter(grid, np.inf), 0.5 * (grid[1:] + grid[:-1])])
    e = np.tile(e, 2)
    imc = np.repeat(np.arange(2, dtype=np.int32), len(e) // 2)
    le = np.log(e)
    for pt in (0, 1):
        _print(f"lookup at grid nodes ptype={pt}: {len(e)} queries")
        assert same_bits(gpu_lookup(ctx1, pt, imc, e, le, dev), oracle.lookup(tables1, pt, imc, e, le))


def test_lookup_invalid_couple_gives_nan(ctx1, dev):
    out = gpu_lookup(ctx1, 0, np.array([5, -1, 0], np.int32), np.array([1.0, 1.0, 1.0]), np.zeros(3), dev)
    assert np.all(np.isnan(out[:, :2])) and np.all(np.isfinite(out[:, 2]))


# ---- config-2 steps -------------------------------------------------------------------------------

@pytest.mark.parametrize("r", [0, 1, 2])
def test_step_golden(golden, ctx2, dev, r):
    _print(f"step golden regime {r}")
    tin = golden_tracks(golden, f"st{r}_in")
    t, res, nrej = gpu_step(ctx2, STEP_SEED, tin, dev)
    assert nrej == 0
    ref_t = {k: golden[f"st{r}_out_{k}"] for k in TRACK_FIELDS}
    ref_r = {k: golden[f"st{r}_res_{k}"] for k in OUT_FIELDS}
    frac = compare_step(tin, t, res, ref_t, ref_r, label=f"regime{r}")
    _print(f"  bit-exact fraction of float outputs: {frac:.6f}")


def test_step_chain_golden(golden, ctx2, dev):
    cur = golden_tracks(golden, "ch_in")
    for s in range(4):
        _print(f"step chain {s}")
        alive = cur["ekin"] > 0
        nxt, res, nrej = gpu_step(ctx2, STEP_SEED + 1, cur, dev)
        assert nrej == (~alive).sum()
        ref_r = {k: golden[f"ch{s}_res_{k}"] for k in OUT_FIELDS}
        ref_t = dict(nxt)
        for k in ("ekin", "dir", "num_ia_left", "rng_ctr", "rng_gauss", "rng_has_gauss", "msc_init_range",
                  "msc_dyn_range_factor", "msc_tlimit_min", "msc_first_step"):
            ref_t[k] = golden[f"ch{s}_out_{k}"]
        compare_step(cur, nxt, res, ref_t, ref_r, alive=alive, label=f"chain{s}")
        # continue from the reference state so each step is checked on its own
        cur = dict(nxt)
        for k in ref_t:
            cur[k] = ref_t[k]


def test_step_stream_high_word_golden(golden, ctx2, dev):
    _print("step golden with 64-bit stream selectors")
    tin = golden_tracks(golden, "hs_in")
    t, res, nrej = gpu_step(ctx2, STEP_SEED + 3, tin, dev)
    assert nrej == 0
    ref_t = {k: golden[f"hs_out_{k}"] for k in TRACK_FIELDS}
    ref_r = {k: golden[f"hs_res_{k}"] for k in OUT_FIELDS}
    compare_step(tin, t, res, ref_t, ref_r, label="stream-hi")


def test_step_parity_1M(tables2, ctx2, oracle, dev):
    """The full config-2 shape: 1M tracks, e-:e+:gamma = 4:1:5, 100 keV-10 GeV."""
    from g4hepem_amd.tracks import step_tracks
    from test_oracle import check_step_invariants
    n = 1_000_000
    tr = step_tracks(n, 20260, 4)
    _print("step parity 1M: gpu")
    t, res, nrej = gpu_step(ctx2, 0xC0FFEE, tr, dev)
    assert nrej == 0
    _print("step parity 1M: oracle")
    ref_t, ref_r = oracle.step(tables2, 0xC0FFEE, tr, nthreads=16)
    frac = compare_step(tr, t, res, ref_t, ref_r, label="1M")
    _print(f"  bit-exact fraction of float outputs: {frac:.6f}")
    check_step_invariants(tr, t, res)


def test_step_wide_energy_parity(tables2, ctx2, oracle, dev):
    from g4hepem_amd.tracks import step_tracks
    for k, (lo, hi) in enumerate(((1e-3, 1.0), (1e3, 1e8))):
        _print(f"wide energy parity {lo}-{hi}")
        tr = step_tracks(200000, 300 + k, 4, lnemin=math.log(lo), lnemax=math.log(hi))
        tr["on_boundary"][:] = (np.arange(200000) % 7 == 3).astype(np.uint8)
        t, res, _ = gpu_step(ctx2, 99, tr, dev)
        ref_t, ref_r = oracle.step(tables2, 99, tr, nthreads=16)
        compare_step(tr, t, res, ref_t, ref_r, label=f"wide{k}")


@pytest.mark.parametrize("form", ["fp64", "fp32"])
def test_process_sort_equals_fused(tables2, dev, form):
    """The two-stage step (particle sort, then process sort of the pending
    interactions) equals the single fused kernel bit for bit."""
    from g4hepem_amd import HEPEM_FP32T, HEPEM_FP64, Context
    from g4hepem_amd.tracks import step_tracks
    ctx = Context(tables2, HEPEM_FP64 if form == "fp64" else HEPEM_FP32T)
    tr = step_tracks(300000, 71, 4, lnemin=math.log(1e-3), lnemax=math.log(1e7))
    tr["on_boundary"][:] = (np.arange(300000) % 9 == 4).astype(np.uint8)
    tr["imc"][17] = 9  # one rejected track
    ctx.set_process_sort(True)
    a_t, a_r, a_rej = gpu_step(ctx, 41, tr, dev)
    ctx.set_process_sort(False)
    b_t, b_r, b_rej = gpu_step(ctx, 41, tr, dev)
    assert a_rej == b_rej == 1
    for k in OUT_FIELDS:
        assert same_bits(a_r[k], b_r[k]), k
    el = tr["ptype"] != 2
    for k, (dt, c) in TRACK_FIELDS.items():
        m = np.tile(el, c) if k.startswith("msc") else np.ones(c * len(el), bool)
        assert same_bits(a_t[k][m], b_t[k][m]), k


@pytest.mark.parametrize("form", ["fp64", "fp32"])
@pytest.mark.parametrize("mix", ["4:1:5", "shuffled", "e-/e+ only", "gamma only"])
def test_tile_local_select_covers_every_slot(tables2, dev, mix, form):
    """The tile-local select stage maps blocks to (tile, chunk) per XCD, with
    the last 8 tiles of each XCD's share split into e-/e+ and gamma halves:
    batches smaller than one tile, ragged last tiles, XCDs without tiles,
    shares shorter and longer than the tail, and tiles whose e-/e+ count is
    not a multiple of 64 (a shuffled layout) all step bit for bit like the
    fused kernel, in both forms."""
    from g4hepem_amd import HEPEM_FP32T, HEPEM_FP64, Context
    from g4hepem_amd.tracks import step_tracks
    ctx = Context(tables2, HEPEM_FP64 if form == "fp64" else HEPEM_FP32T)
    for n in (1, 63, 4095, 4097, 8 * 4096 + 1, 40000, 200000, 300001):
        _print(f"tile-local select, {form}, mix {mix}, n {n}")
        tr = step_tracks(n, 500 + n % 97, 4, lnemin=math.log(1e-2), lnemax=math.log(1e5))
        rng = np.random.default_rng(n)
        if mix == "shuffled":
            tr["ptype"][:] = rng.permutation(tr["ptype"])
        elif mix == "e-/e+ only":
            tr["ptype"][:] = np.where(rng.random(n) < 0.8, 0, 1).astype(np.int8)
        elif mix == "gamma only":
            tr["ptype"][:] = 2
        ctx.set_process_sort(True)
        a_t, a_r, _ = gpu_step(ctx, 43, tr, dev)
        ctx.set_process_sort(False)
        b_t, b_r, _ = gpu_step(ctx, 43, tr, dev)
        for k in OUT_FIELDS:
            assert same_bits(a_r[k], b_r[k]), (n, k)
        el = tr["ptype"] != 2
        for k, (dt, c) in TRACK_FIELDS.items():
            m = np.tile(el, c) if k.startswith("msc") else np.ones(c * n, bool)
            assert same_bits(a_t[k][m], b_t[k][m]), (n, k)


@pytest.mark.parametrize("form", ["fp64", "fp32"])
def test_step_prepare_run_across_a_mode_change(tables2, dev, form):
    """hepem_step_prepare lays the queue out for the process-sort mode of its
    call and records the layout in the work buffer; hepem_step_run reads it,
    so a mode changed between the two calls (either way) still steps every
    track as hepem_step does."""
    from g4hepem_amd import HEPEM_FP32T, HEPEM_FP64, Context
    from g4hepem_amd.tracks import StepOutSoA, TrackSoA, step_outputs_differ, step_tracks
    ctx = Context(tables2, HEPEM_FP64 if form == "fp64" else HEPEM_FP32T)
    n = 70001
    tr = step_tracks(n, 88, 4, lnemin=math.log(1e-2), lnemax=math.log(1e5))
    tr["ptype"][:] = np.random.default_rng(88).permutation(tr["ptype"])
    tin = TrackSoA.from_numpy(tr, "torch", dev)
    work = ctx.work_buffer(n, dev)
    ref_t, ref_r = TrackSoA(n, "torch", dev), StepOutSoA(n, "torch", dev)
    ctx.step(13, tin, ref_t, ref_r, work)
    for before, after in ((True, False), (False, True)):
        _print(f"prepare with sort {before}, run with sort {after}, {form}")
        t, r = TrackSoA(n, "torch", dev), StepOutSoA(n, "torch", dev)
        ctx.set_process_sort(before)
        ctx.step_prepare(tin, work)
        ctx.set_process_sort(after)
        ctx.step_run(13, tin, t, r, work)
        torch.cuda.synchronize()
        diff = step_outputs_differ(tin["ptype"], t, r, ref_t, ref_r)
        assert not diff, (before, after, diff)
    ctx.set_process_sort("auto")


def test_step_with_tables_from_json(tables2, ctx2, dev):
    """Tables written as a G4HepEmState JSON document and read back drive the
    step bit for bit like the originals (the device image is rebuilt from
    the parsed tables)."""
    from g4hepem_amd import HEPEM_FP64, Context
    from g4hepem_amd._native import Tables
    from g4hepem_amd.tracks import step_tracks
    ctx = Context(Tables.from_json(tables2.to_json()), HEPEM_FP64)
    tr = step_tracks(50000, 12, 4, lnemin=math.log(1e-3), lnemax=math.log(1e7))
    _print("step with tables from JSON")
    a_t, a_r, _ = gpu_step(ctx2, 9, tr, dev)
    b_t, b_r, _ = gpu_step(ctx, 9, tr, dev)
    for k in OUT_FIELDS:
        assert same_bits(a_r[k], b_r[k]), k
    for k in ("ekin", "dir", "num_ia_left", "rng_ctr"):
        assert same_bits(a_t[k], b_t[k]), k


def test_step_inplace_equals_out_of_place(ctx2, dev):
    from g4hepem_amd.tracks import step_tracks
    tr = step_tracks(50000, 5, 4)
    a_t, a_r, _ = gpu_step(ctx2, 7, tr, dev, inplace=False)
    b_t, b_r, _ = gpu_step(ctx2, 7, tr, dev, inplace=True)
    for k in OUT_FIELDS:
        assert same_bits(a_r[k], b_r[k]), k
    for k in ("ekin", "dir", "num_ia_left", "rng_ctr"):
        assert same_bits(a_t[k], b_t[k]), k


def test_step_rejects_invalid_tracks(ctx2, dev):
    from g4hepem_amd.tracks import step_tracks
    tr = step_tracks(1000, 8, 4)
    tr["imc"][3] = 99
    tr["imc"][4] = -1
    tr["ekin"][5] = 0.0
    tr["ekin"][6] = -1.0
    tr["ekin"][7] = np.nan
    t, res, nrej = gpu_step(ctx2, 1, tr, dev)
    assert nrej == 5
    assert list(res["winner"][3:8]) == [-99] * 5
    assert np.all(res["winner"][8:] != -99)
    assert same_bits(t["ekin"][3:8], tr["ekin"][3:8])


def test_reset_state(ctx2, dev):
    from g4hepem_amd import lib
    from g4hepem_amd.tracks import TrackSoA, step_tracks
    tr = step_tracks(777, 3, 4)
    t = TrackSoA.from_numpy(tr, "torch", dev)
    t["num_ia_left"].fill_(5.0)
    t["rng_ctr"].fill_(9)
    s = t.struct()
    import ctypes as C
    assert lib().hepem_tracks_reset_state(C.byref(s), None) == 0
    torch.cuda.synchronize()
    a = t.to_numpy()
    for k in ("num_ia_left", "msc_init_range", "msc_dyn_range_factor", "msc_tlimit_min", "msc_first_step", "rng_ctr"):
        assert same_bits(a[k], tr[k]), k


# ---- FP32-table form ------------------------------------------------------------------------------

def test_fp32_lookup_errors(tables1, oracle, dev):
    """FP32-table lookups against the FP64 oracle: report max/mean relative error."""
    from g4hepem_amd import HEPEM_FP32T, Context
    from g4hepem_amd.tracks import lookup_queries
    ctx = Context(tables1, HEPEM_FP32T)
    imc, e, le = lookup_queries(100000, 11, 2)
    names = ("dedx", "range", "invrange", "xs_ioni", "xs_brem")
    for pt in (0, 1):
        ref = oracle.lookup(tables1, pt, imc, e, le)
        out = gpu_lookup(ctx, pt, imc, e, le, dev)
        for q in range(5):
            m = ref[q] > 0
            rel = np.abs(out[q][m] - ref[q][m]) / ref[q][m]
            _print(f"  fp32 ptype={pt} {names[q]}: max rel {rel.max():.3e} mean rel {rel.mean():.3e}")
            assert rel.max() < 1e-5, names[q]


def test_fp32_step_statistics(tables2, ctx2_32, oracle, dev):
    """FP32-table step: integer decisions follow the FP64 oracle for the vast
    majority of tracks and conservation holds at FP32 precision."""
    from g4hepem_amd.tracks import step_tracks
    from test_oracle import check_step_invariants
    n = 200000
    tr = step_tracks(n, 44, 4)
    t, res, nrej = gpu_step(ctx2_32, 5, tr, dev)
    assert nrej == 0
    ref_t, ref_r = oracle.step(tables2, 5, tr, nthreads=16)
    agree = (res["winner"] == ref_r["winner"]).mean()
    _print(f"  fp32 winner agreement {agree:.5f}")
    assert agree > 0.99
    # energy deposited by all e-/e+ (continuous loss + fluctuation) agrees with FP64
    el = tr["ptype"] != 2
    rel = abs(res["edep"][el].sum() - ref_r["edep"][el].sum()) / ref_r["edep"][el].sum()
    _print(f"  fp32 total e-/e+ deposit rel. difference {rel:.2e}")
    assert rel < 1e-3
    check_step_invariants(tr, t, res, tol=2e-5, dir_tol=2e-6)
