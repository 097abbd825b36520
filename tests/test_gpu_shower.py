"""Showers in the 50 x (2.3 mm Pb + 5.7 mm lAr) slab stack (hepem_slab_run)
against the oracle's CPU harness of the same tracking loop, event by event:
deposits are per-track deterministic (per-track streams, integer scoring), so
most events agree exactly and the per-slab means agree to statistics (chi2)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

LAYERS, THICK = 50, (2.3, 5.7)


def _print(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def ctx1(tables1):
    assert torch.cuda.is_available(), "GPU test run without a visible GPU"
    from g4hepem_amd import HEPEM_FP64, Context
    return Context(tables1, HEPEM_FP64)


def chi2_means(a, b):
    """chi2 of the per-slab means of two event samples (rows = events)."""
    ma, mb = a.mean(0), b.mean(0)
    va, vb = a.var(0, ddof=1) / len(a), b.var(0, ddof=1) / len(b)
    v = va + vb
    m = v > 0
    return float((((ma - mb) ** 2)[m] / v[m]).sum()), int(m.sum())


def test_shower_matches_cpu_harness(tables1, ctx1, oracle):
    from g4hepem_amd._native import slab_setup
    nev, e0 = 200, 1000.0
    su = slab_setup(LAYERS, THICK, [0, 1], 0, e0, capacity=1 << 21, events_per_chunk=64)
    _print("shower 1 GeV e-: gpu")
    g = ctx1.slab_run(su, nev, 77, per_event=True)
    _print(f"  steps {g['steps']} tracks {g['tracks']} iterations {g['iterations']} max in flight {g['max_in_flight']}")
    _print("shower 1 GeV e-: cpu harness")
    dep, leak, steps, tracks = oracle.slab_shower(tables1, su, nev, 77, nthreads=16)
    # energy balance of every event on both sides (fixed-point rounding only)
    for d, lk in ((g["edep_event"], g["leak_event"]), (dep, leak)):
        bal = (d.sum(1) + lk) / 1e9
        assert np.abs(bal - e0).max() < 1e-4
    same = np.all(g["edep_event"] == dep, axis=1) & (g["leak_event"] == leak)
    _print(f"  events identical to the CPU harness: {same.mean():.3f}")
    assert same.mean() > 0.5
    chi2, ndf = chi2_means(g["edep_event"] / 1e9, dep / 1e9)
    _print(f"  chi2/ndf of per-slab means: {chi2:.1f}/{ndf}")
    assert chi2 < ndf + 5 * np.sqrt(2 * ndf)
    # sums reported by the run equal the per-event records
    assert np.allclose(g["edep_sum"], (g["edep_event"] / 1e9).sum(0), rtol=1e-12, atol=1e-9)
    assert abs(g["leak_sum"] - g["leak_event"].sum() / 1e9) < 1e-6
    assert abs(steps - g["steps"]) < 0.02 * steps


def paired_chi2(a, b):
    """chi2 of the per-slab mean differences of paired events (rows = events,
    the same seeds and streams on both sides)."""
    d = a - b
    se2 = d.var(0, ddof=1) / len(d)
    m = se2 > 0
    return float(((d.mean(0) ** 2)[m] / se2[m]).sum()), int(m.sum())


@pytest.mark.skipif(not __import__("oracle.oracle_py", fromlist=["ref_available"]).ref_available(),
                    reason="compiled reference (oracle/_ref) not shipped")
def test_shower_10gev_matches_reference_harness(tables1, ctx1):
    """200 events of the benchmark's 10 GeV e- showers against the same
    showers stepped by the reference's own compiled code (TrackElectron's MSC
    sub-step loop, TrackGamma, StackSecondaries: oracle/ref/ref_harness.cc)."""
    from g4hepem_amd._native import slab_setup
    from oracle.oracle_py import Ref
    nev, e0 = 200, 1.0e4
    su = slab_setup(LAYERS, THICK, [0, 1], 0, e0, capacity=1 << 23, events_per_chunk=200)
    _print("shower 10 GeV e-: gpu")
    g = ctx1.slab_run(su, nev, 31, per_event=True)
    _print("shower 10 GeV e-: reference harness")
    dep, leak, steps, tracks = Ref().slab_shower(tables1, su, nev, 31, nthreads=16)
    for d, lk in ((g["edep_event"], g["leak_event"]), (dep, leak)):
        assert np.abs((d.sum(1) + lk) / 1e9 - e0).max() < 1e-4
    same = np.all(g["edep_event"] == dep, axis=1) & (g["leak_event"] == leak)
    chi2, ndf = paired_chi2(g["edep_event"] / 1e9, dep / 1e9)
    _print(f"  identical events {same.mean():.3f}, paired chi2/ndf {chi2:.1f}/{ndf}, steps {g['steps']} vs {steps}")
    assert same.mean() > 0.2
    assert chi2 < ndf + 5 * np.sqrt(2 * ndf)
    assert abs(steps - g["steps"]) < 0.01 * steps and abs(tracks - g["tracks"]) < 0.01 * tracks


def test_shower_substeps_off_matches_cpu_harness(tables1, oracle):
    """With fIsMultipleStepsInMSCTrans off in every region (through the JSON
    form) every e-/e+ step is one HowFar/Perform pass; GPU and CPU harness agree
    on those tables too, and take more steps than with sub-steps."""
    import json
    from g4hepem_amd import HEPEM_FP64, Context
    from g4hepem_amd._native import Tables, slab_setup
    doc = json.loads(tables1.to_json())
    for r in doc["fParameters"]["fParametersPerRegion"]:
        r["fIsMultipleStepsInMSCTrans"] = False
    t_off = Tables.from_json(json.dumps(doc))
    nev, e0 = 100, 1000.0
    su = slab_setup(LAYERS, THICK, [0, 1], 0, e0, capacity=1 << 21, events_per_chunk=64)
    g = Context(t_off, HEPEM_FP64).slab_run(su, nev, 8, per_event=True)
    dep, leak, steps, tracks = oracle.slab_shower(t_off, su, nev, 8, nthreads=16)
    same = np.all(g["edep_event"] == dep, axis=1) & (g["leak_event"] == leak)
    _print(f"  sub-steps off: identical events {same.mean():.3f}, steps {g['steps']} vs {steps}")
    assert same.mean() > 0.5
    on = Context(tables1, HEPEM_FP64).slab_run(su, nev, 8)
    assert on["steps"] < g["steps"]


@pytest.mark.parametrize("chunk,refill", [(7, 0), (3, 2000), (5, 1 << 30)])
def test_shower_independent_of_chunking(ctx1, chunk, refill):
    """Per-event results do not depend on how many events are stacked together
    or on when the next events join the tracks still in flight (refill 1<<30:
    new primaries whenever the first events' growth period has passed)."""
    from g4hepem_amd._native import slab_setup
    a = ctx1.slab_run(slab_setup(LAYERS, THICK, [0, 1], 2, 500.0, capacity=1 << 20, events_per_chunk=chunk,
                                 refill_below=refill), 40, 5, per_event=True)
    b = ctx1.slab_run(slab_setup(LAYERS, THICK, [0, 1], 2, 500.0, capacity=1 << 20, events_per_chunk=40), 40, 5,
                      per_event=True)
    assert np.array_equal(a["edep_event"], b["edep_event"]) and np.array_equal(a["leak_event"], b["leak_event"])
    assert a["steps"] == b["steps"] and a["tracks"] == b["tracks"]
    assert a["edep_sum"].tobytes() == b["edep_sum"].tobytes()


def test_shower_capacity_error(ctx1):
    from g4hepem_amd import HepEmError
    from g4hepem_amd._native import slab_setup
    with pytest.raises(HepEmError):
        ctx1.slab_run(slab_setup(LAYERS, THICK, [0, 1], 0, 10000.0, capacity=64, events_per_chunk=4), 4, 1)


def test_shower_photon_transmission_closed_form(tables1, ctx1, oracle):
    """1 MeV photons on one 2 mm Pb slab: the fraction of events without any
    deposit is the no-interaction probability exp(-d Sigma_tot)."""
    import math
    from g4hepem_amd._native import slab_setup
    d, e0, nev = 2.0, 1.0, 50000
    g = ctx1.slab_run(slab_setup(1, (d,), (0,), 2, e0, capacity=1 << 20, events_per_chunk=50000), nev, 9, per_event=True)
    zero = (g["edep_event"].sum(1) == 0).mean()
    xs, _, _ = oracle.probe_gamma(tables1, np.array([0], np.int32), np.array([e0]), np.array([math.log(e0)]),
                                  np.array([0.5]))
    p = math.exp(-d * xs[0])
    assert abs(zero - p) < 4 * math.sqrt(p * (1 - p) / nev), (zero, p)


def test_shower_production_cuts_follow_region_flag(tables1, ctx1, oracle):
    """Secondaries below the couple's production threshold are deposited, not
    stacked, exactly where the region applies cuts (StackSecondaries,
    G4HepEmTrackingManager.cc:1274-1313). The same tables with fIsApplyCuts
    switched off (through the JSON form) stack every secondary; both settings
    agree with the CPU harness and conserve every event's energy."""
    import json
    from g4hepem_amd import HEPEM_FP64, Context
    from g4hepem_amd._native import Tables, slab_setup
    doc = json.loads(tables1.to_json())
    assert all(r["fIsApplyCuts"] for r in doc["fParameters"]["fParametersPerRegion"])
    for r in doc["fParameters"]["fParametersPerRegion"]:
        r["fIsApplyCuts"] = False
    nocut = Tables.from_json(json.dumps(doc))
    nev, e0 = 60, 1000.0
    su = slab_setup(LAYERS, THICK, [0, 1], 0, e0, capacity=1 << 21, events_per_chunk=30)
    runs = {}
    for name, tables, ctx in (("cuts", tables1, ctx1), ("no cuts", nocut, Context(nocut, HEPEM_FP64))):
        _print(f"shower with {name}")
        g = ctx.slab_run(su, nev, 78, per_event=True)
        dep, leak, steps, tracks = oracle.slab_shower(tables, su, nev, 78, nthreads=16)
        bal = (g["edep_event"].sum(1) + g["leak_event"]) / 1e9
        assert np.abs(bal - e0).max() < 1e-4
        same = np.all(g["edep_event"] == dep, axis=1) & (g["leak_event"] == leak)
        _print(f"  tracks {g['tracks']} (cpu {tracks}), identical events {same.mean():.3f}")
        assert same.mean() > 0.5
        runs[name] = g
    assert runs["cuts"]["tracks"] < 0.8 * runs["no cuts"]["tracks"]


def test_shower_woodcock_matches_cpu_harness(tables1, ctx1, oracle):
    """Woodcock tracking of the gammas above 0.2 MeV through the whole stack
    (G4HepEmWoodcockHelper.cc:145-300): the device driver against the CPU
    harness, which runs the same delta steps between its HowFar and Perform,
    event by event."""
    from g4hepem_amd._native import slab_setup
    nev, e0 = 200, 1000.0
    su = slab_setup(LAYERS, THICK, [0, 1], 0, e0, capacity=1 << 21, events_per_chunk=64, woodcock_min_ekin=0.2)
    _print("shower 1 GeV e-, Woodcock gammas: gpu")
    g = ctx1.slab_run(su, nev, 79, per_event=True)
    _print(f"  steps {g['steps']} tracks {g['tracks']} iterations {g['iterations']}")
    _print("shower 1 GeV e-, Woodcock gammas: cpu harness")
    dep, leak, steps, tracks = oracle.slab_shower(tables1, su, nev, 79, nthreads=16)
    for d, lk in ((g["edep_event"], g["leak_event"]), (dep, leak)):
        assert np.abs((d.sum(1) + lk) / 1e9 - e0).max() < 1e-4
    same = np.all(g["edep_event"] == dep, axis=1) & (g["leak_event"] == leak)
    _print(f"  events identical to the CPU harness: {same.mean():.3f}")
    assert same.mean() > 0.5
    chi2, ndf = chi2_means(g["edep_event"] / 1e9, dep / 1e9)
    _print(f"  chi2/ndf of per-slab means: {chi2:.1f}/{ndf}")
    assert chi2 < ndf + 5 * np.sqrt(2 * ndf)
    assert abs(steps - g["steps"]) < 0.02 * steps


def test_shower_woodcock_transmission_closed_form(tables1, ctx1, oracle):
    """1 MeV photons through 5.7 mm lAr then 2 mm Pb with Woodcock tracking:
    the fraction of events without any deposit is
    exp(-(5.7 Sigma_lAr + (2 - 1e-3) Sigma_Pb)) (the gamma leaves the
    Woodcock envelope 1e-3 mm before its exit)."""
    import math
    from g4hepem_amd._native import slab_setup
    e0, nev = 1.0, 50000
    su = slab_setup(1, (5.7, 2.0), (1, 0), 2, e0, capacity=1 << 20, events_per_chunk=50000, woodcock_min_ekin=0.2)
    g = ctx1.slab_run(su, nev, 10, per_event=True)
    zero = (g["edep_event"].sum(1) == 0).mean()
    xs, _, _ = oracle.probe_gamma(tables1, np.array([1, 0], np.int32), np.array([e0, e0]),
                                  np.array([math.log(e0)] * 2), np.array([0.5, 0.5]))
    p = math.exp(-(5.7 * xs[0] + (2.0 - 1.0e-3) * xs[1]))
    assert abs(zero - p) < 4 * math.sqrt(p * (1 - p) / nev), (zero, p)


@pytest.mark.parametrize("form", ["fp64", "fp32t"])
def test_shower_woodcock_agrees_with_normal_tracking(tables1, ctx1, form):
    """1 GeV e- showers: with Woodcock tracking of the gammas the per-slab
    mean deposits agree with normal tracking (chi2), every event conserves
    energy, fewer steps are taken, and the run does not depend on chunking."""
    from g4hepem_amd import HEPEM_FP32T, Context
    from g4hepem_amd._native import slab_setup
    ctx = ctx1 if form == "fp64" else Context(tables1, HEPEM_FP32T)
    nev, e0 = 1000, 1000.0
    runs = {}
    for w in (0.0, 0.2):
        su = slab_setup(LAYERS, THICK, [0, 1], 0, e0, capacity=1 << 22, events_per_chunk=250, woodcock_min_ekin=w)
        g = ctx.slab_run(su, nev, 80, per_event=True)
        assert np.abs((g["edep_event"].sum(1) + g["leak_event"]) / 1e9 - e0).max() < 1e-4
        runs[w] = g
        _print(f"  {form} woodcock {w}: steps {g['steps']}")
    chi2, ndf = chi2_means(runs[0.0]["edep_event"] / 1e9, runs[0.2]["edep_event"] / 1e9)
    _print(f"  chi2/ndf of per-slab means: {chi2:.1f}/{ndf}")
    assert chi2 < ndf + 5 * np.sqrt(2 * ndf)
    assert runs[0.2]["steps"] < 0.8 * runs[0.0]["steps"]
    su = slab_setup(LAYERS, THICK, [0, 1], 0, e0, capacity=1 << 22, events_per_chunk=7, refill_below=3000,
                    woodcock_min_ekin=0.2)
    c = ctx.slab_run(su, 100, 80, per_event=True)
    assert np.array_equal(c["edep_event"], runs[0.2]["edep_event"][:100])


def test_shower_setup_errors_and_empty_run(ctx1):
    """Invalid set-ups are rejected before any work; a run of no events
    returns zero sums."""
    import math
    from g4hepem_amd import HepEmError
    from g4hepem_amd._native import slab_setup
    for w in (-1.0, math.nan):
        with pytest.raises(HepEmError):
            ctx1.slab_run(slab_setup(LAYERS, THICK, [0, 1], 0, 100.0, capacity=1024, woodcock_min_ekin=w), 2, 1)
    with pytest.raises(HepEmError):
        ctx1.slab_run(slab_setup(LAYERS, THICK, [0, 99], 0, 100.0, capacity=1024), 2, 1)
    g = ctx1.slab_run(slab_setup(LAYERS, THICK, [0, 1], 0, 100.0, capacity=1024, woodcock_min_ekin=0.2), 0, 1,
                      per_event=True)
    assert g["steps"] == 0 and g["tracks"] == 0 and not g["edep_sum"].any() and g["leak_sum"] == 0.0


@pytest.mark.parametrize("woodcock", [0.0, 0.2])
def test_shower_independent_of_process_sort(tables1, woodcock):
    """The driver's Perform calls give the same events with the process sort
    on, off or automatic (the sort only changes which thread samples a slot)."""
    from g4hepem_amd import HEPEM_FP64, Context
    from g4hepem_amd._native import slab_setup
    su = slab_setup(LAYERS, THICK, [0, 1], 0, 1000.0, capacity=1 << 21, events_per_chunk=40, woodcock_min_ekin=woodcock)
    ctx = Context(tables1, HEPEM_FP64)
    runs = []
    for mode in ("auto", True, False):
        ctx.set_process_sort(mode)
        runs.append(ctx.slab_run(su, 40, 81, per_event=True))
    for g in runs[1:]:
        assert np.array_equal(g["edep_event"], runs[0]["edep_event"])
        assert np.array_equal(g["leak_event"], runs[0]["leak_event"])
        assert g["steps"] == runs[0]["steps"] and g["tracks"] == runs[0]["tracks"]
