"""The device track stack (hepem_stack_*) through the C-ABI: a batch pushed,
stepped in place, and advanced into the next batch, against StackSecondaries'
rules (G4HepEm/G4HepEm/src/G4HepEmTrackingManager.cc:1254-1330) restated
per parent in numpy below. That function needs Geant4 and is not compiled
here: its rules are pinned by the slab shower, whose driver applies them and
matches the harness over the reference's own functions event by event.

Checked bit for bit: every field of the next batch (survivors as they left
the step, stacked secondaries with their parent's couple, safety and boundary
flag, start-of-track state and derived streams), the e-/e+-then-gamma order,
parent and origin, and the energy the production cuts deposit; with cuts
applied (the e+ rule exercised by a couple whose gamma threshold is above the
electron mass) and without; over two passes; and the capacity error leaving
the batch unchanged.
"""
import json
import math

import numpy as np
import pytest

from conftest import SET2, same_bits
from oracle.oracle_py import TRACK_FIELDS

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

ME = 0.51099895000
M64 = (1 << 64) - 1


def _print(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "GPU test run without a visible GPU"
    return torch.device("cuda:0")


def stack_tables(apply: bool):
    """SET2 at 0.7 mm with the regions' fIsApplyCuts set to `apply`; couple 3
    gets a gamma threshold of 0.6 MeV (above m_e c^2), so its e+ secondaries
    below the e+ threshold are deposited."""
    from g4hepem_amd import Tables
    doc = json.loads(Tables(SET2, 0.7).to_json())
    for r in doc["fParameters"]["fParametersPerRegion"]:
        r["fIsApplyCuts"] = apply
    doc["fData"]["fTheMatCutData"]["fMatCutData"][3]["fSecGamProdCutE"] = 0.6
    cuts = [(m["fSecElProdCutE"], m["fSecPosProdCutE"], m["fSecGamProdCutE"],
             doc["fParameters"]["fParametersPerRegion"][m["fG4RegionIndex"]]["fIsApplyCuts"])
            for m in doc["fData"]["fTheMatCutData"]["fMatCutData"]]
    return Tables.from_json(json.dumps(doc)), cuts


def mix64(x):
    x = (x + 0x9E3779B97F4A7C15) & M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M64
    return x ^ (x >> 31)


def expected_next(t, r, keep, cuts):
    """StackSecondaries per parent, in parent order: the track if kept and
    alive, then each secondary slot, deposited below threshold where the
    couple's region applies cuts. Returns the next batch (e-/e+ then
    gammas), parent, origin and the cut deposit per parent."""
    n = len(t["ekin"])
    rowsE, rowsG, cut = [], [], np.zeros(n)
    for i in range(n):
        if keep[i] and t["ekin"][i] > 0.0:
            (rowsG if t["ptype"][i] == 2 else rowsE).append((i, 0))
        ce, cp, cg, apply = cuts[t["imc"][i]]
        for s in range(int(r["nsec"][i])):
            pt, se = int(r["sec_ptype"][s * n + i]), float(r["sec_ekin"][s * n + i])
            c = apply and ((pt == 0 and se < ce) or (pt == 1 and ME < cg and se < cp) or (pt == 2 and se < cg))
            if c:
                cut[i] += se + 2.0 * ME if pt == 1 else se
            else:
                (rowsG if pt == 2 else rowsE).append((i, s + 1))
    rows = rowsE + rowsG
    m = len(rows)
    nxt = {k: np.zeros(c * m, dtype=np.dtype(dt)) for k, (dt, c) in TRACK_FIELDS.items()}
    for j, (i, o) in enumerate(rows):
        if o == 0:
            for k, (_, c) in TRACK_FIELDS.items():
                for q in range(c):
                    nxt[k][q * m + j] = t[k][q * n + i]
            continue
        s = o - 1
        sid = (int(t["rng_stream"][i]) << 32) | int(np.uint32(t["id"][i]))
        cid = mix64(sid ^ mix64((int(t["rng_ctr"][i]) << 2) | (s + 1)))
        nxt["ptype"][j] = r["sec_ptype"][s * n + i]
        nxt["id"][j] = np.uint32(cid & 0xFFFFFFFF).view(np.int32)
        nxt["rng_stream"][j] = cid >> 32
        nxt["imc"][j] = t["imc"][i]
        nxt["ekin"][j] = r["sec_ekin"][s * n + i]
        for c in range(3):
            nxt["dir"][c * m + j] = r["sec_dir"][(c * 2 + s) * n + i]
        nxt["safety"][j] = t["safety"][i]
        nxt["on_boundary"][j] = t["on_boundary"][i]
        for q in range(4):
            nxt["num_ia_left"][q * m + j] = -1.0
        nxt["msc_init_range"][j] = 1.0e21
        nxt["msc_dyn_range_factor"][j] = 0.04
        nxt["msc_tlimit_min"][j] = 1.0e-7
        nxt["msc_first_step"][j] = 1
    parent = np.array([i for i, _ in rows], dtype=np.int32)
    origin = np.array([o for _, o in rows], dtype=np.int8)
    return nxt, parent, origin, cut, len(rowsE)


def step_and_advance(ctx, stack, seed, keep_frac, dev, with_outputs=True):
    from g4hepem_amd.tracks import StepOutSoA, TrackSoA
    v = stack.view()
    n = v.n
    res = StepOutSoA(n, "torch", dev)
    ctx.step(seed, v, v, res)    # in place, on the stack's own buffer
    post = stack.copy_out(TrackSoA(n, "torch", dev)).to_numpy()
    r = res.to_numpy()
    g = torch.Generator(device="cpu").manual_seed(seed)
    keep_h = (torch.rand(n, generator=g) < keep_frac).to(torch.uint8)
    keep = keep_h.to(dev)
    cap = 3 * n      # a track and its two secondaries at most
    cut = torch.full((n,), -1.0, dtype=torch.float64, device=dev)
    parent = torch.full((cap,), -7, dtype=torch.int32, device=dev)
    origin = torch.full((cap,), -7, dtype=torch.int8, device=dev)
    if with_outputs:
        m = stack.advance(v, res, keep, cut, parent, origin)
    else:
        m = stack.advance(v, res, keep)
    torch.cuda.synchronize()
    return post, r, keep_h.numpy(), m, cut.cpu().numpy(), parent.cpu().numpy()[:m], origin.cpu().numpy()[:m]


@pytest.mark.parametrize("apply", [True, False], ids=["cuts", "no-cuts"])
def test_stack_advance_follows_stack_secondaries(dev, apply):
    from g4hepem_amd import HEPEM_FP64, Context, TrackStack
    from g4hepem_amd.tracks import TrackSoA, step_tracks
    tables, cuts = stack_tables(apply)
    ctx = Context(tables, HEPEM_FP64)
    n = 60_000
    tr = step_tracks(n, 41, len(SET2), lnemin=math.log(1.0), lnemax=math.log(1.0e4))
    stack = TrackStack(ctx, 4 * n)
    stack.push(TrackSoA.from_numpy(tr, "torch", dev))
    assert len(stack) == n
    _print(f"stack, apply cuts {apply}: first pass over {n} tracks")
    post, r, keep, m, cut, parent, origin = step_and_advance(ctx, stack, 7, 0.9, dev)
    nxt, eparent, eorigin, ecut, ne = expected_next(post, r, keep, cuts)
    assert m == len(eparent)
    got = stack.copy_out(TrackSoA(m, "torch", dev)).to_numpy()
    for k in TRACK_FIELDS:
        assert same_bits(got[k], nxt[k]), k
    assert same_bits(parent, eparent) and same_bits(origin, eorigin)
    assert same_bits(cut, ecut)
    assert (got["ptype"][:ne] != 2).all() and (got["ptype"][ne:] == 2).all()
    nsec = r["nsec"].sum()
    stacked = (origin > 0).sum()
    _print(f"  next batch {m}: {ne} e-/e+, {(origin == 0).sum()} survivors, {stacked} of {nsec} secondaries stacked, "
           f"cut deposit {cut.sum():.3f} MeV")
    if apply:
        assert stacked < nsec and ecut.sum() > 0.0
        # the e+ rule fires on the couple whose gamma threshold exceeds m_e c^2
        epos = np.zeros(len(post["ekin"]), bool)
        nn = len(post["ekin"])
        for s in range(2):
            pt = r["sec_ptype"][s * nn: (s + 1) * nn]
            se = r["sec_ekin"][s * nn: (s + 1) * nn]
            epos |= (s < r["nsec"]) & (pt == 1) & (post["imc"] == 3) & (se < cuts[3][1])
        assert epos.sum() > 0
    else:
        assert stacked == nsec and (cut == 0.0).all()
    # secondaries' streams are distinct from each other and from the tracks'
    sid = (got["rng_stream"].astype(np.uint64) << np.uint64(32)) | got["id"].view(np.uint32).astype(np.uint64)
    assert len(np.unique(sid)) == m
    # a second pass over the next batch, without the optional outputs
    post2, r2, keep2, m2, _, _, _ = step_and_advance(ctx, stack, 8, 0.8, dev, with_outputs=False)
    _, ep2, _, _, _ = expected_next(post2, r2, keep2, cuts)
    assert m2 == len(ep2) == len(stack)
    stack.close()
    ctx.close()


def test_stack_capacity_error_keeps_the_batch(dev):
    from g4hepem_amd import HEPEM_FP64, Context, HepEmError, TrackStack
    from g4hepem_amd.tracks import StepOutSoA, TrackSoA, step_tracks
    tables, _ = stack_tables(False)
    ctx = Context(tables, HEPEM_FP64)
    n = 20_000
    tr = step_tracks(n, 43, len(SET2), lnemin=math.log(100.0), lnemax=math.log(1.0e4))
    stack = TrackStack(ctx, n + 10)     # survivors plus secondaries overflow it
    stack.push(TrackSoA.from_numpy(tr, "torch", dev))
    with pytest.raises(HepEmError) as e:
        stack.push(TrackSoA.from_numpy(tr, "torch", dev))
    assert e.value.status == -3 and len(stack) == n
    v = stack.view()
    res = StepOutSoA(n, "torch", dev)
    ctx.step(3, v, v, res)
    before = stack.copy_out(TrackSoA(n, "torch", dev)).to_numpy()
    with pytest.raises(HepEmError) as e:
        stack.advance(v, res)
    assert e.value.status == -3 and len(stack) == n
    after = stack.copy_out(TrackSoA(n, "torch", dev)).to_numpy()
    for k in TRACK_FIELDS:
        assert same_bits(after[k], before[k]), k
    stack.close()
    ctx.close()
