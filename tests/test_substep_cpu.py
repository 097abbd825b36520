"""The e-/e+ step as G4HepEmTrackingManager::TrackElectron takes it
(G4HepEm/G4HepEm/src/G4HepEmTrackingManager.cc:408-635): HowFarToDiscreteInteraction
once, then MSC-limited sub-steps -- HowFarToMSC, the move, UpdatePStepLength,
UpdateNumIALeft, ApplyMeanEnergyLoss, SampleMSC with its displacement -- while the
region sets fIsMultipleStepsInMSCTrans, then the loss fluctuation and the discrete
interaction. The oracle's restatement (oracle_electron_slab_step, and the slab
shower built on it) against the same loop driving the reference's own compiled
G4HepEmElectronManager functions (oracle/ref/ref_harness.cc), bit for bit, with
sub-stepping on and off; and, with it off and no way out of the world, against
the manager's HowFar + Perform with the same geometry cut.
"""
import json
import math

import numpy as np
import pytest

from conftest import CUT, SET1, SET2, same_bits
from oracle.oracle_py import OUT_FIELDS, TRACK_FIELDS, ref_available

pytestmark = pytest.mark.skipif(not ref_available(), reason="compiled reference (oracle/_ref) not built")


@pytest.fixture(scope="module")
def ref():
    from oracle.oracle_py import Ref
    return Ref()


def tables_multistep(base, on: bool):
    """`base` tables with every region's fIsMultipleStepsInMSCTrans set to `on`
    (through the JSON form)."""
    from g4hepem_amd import Tables
    doc = json.loads(base.to_json())
    for r in doc["fParameters"]["fParametersPerRegion"]:
        r["fIsMultipleStepsInMSCTrans"] = on
    return Tables.from_json(json.dumps(doc))


def slab_tracks(n, seed, ncouples, lnemin=math.log(1e-3), lnemax=math.log(1e4), exits=True):
    """e-/e+ (4:1) with log-uniform energies in slabs of 0.3-50 mm; a fifth on
    one of the planes (on a boundary); bit 0/1 of exitf: leaving through lo/hi
    leaves the world."""
    from g4hepem_amd.tracks import step_tracks
    tr = step_tracks(n, seed, ncouples, lnemin=lnemin, lnemax=lnemax, mix=(4, 1, 0))
    rng = np.random.Generator(np.random.Philox(seed + 1))
    width = np.array([0.3, 2.3, 5.7, 50.0])[rng.integers(0, 4, n)]
    x = rng.random(n) * width
    onb = rng.random(n) < 0.2
    x[onb] = np.where(rng.random(onb.sum()) < 0.5, 0.0, width[onb])
    tr["on_boundary"][:] = onb
    exitf = rng.integers(0, 4, n).astype(np.uint8) if exits else np.zeros(n, np.uint8)
    return tr, x, np.zeros(n), width, exitf


def assert_same(a, b):
    (ta, ra, ga), (tb, rb, gb) = a, b
    for k in TRACK_FIELDS:
        assert same_bits(ta[k], tb[k]), f"state {k}"
    for k in OUT_FIELDS:
        assert same_bits(ra[k], rb[k]), f"result {k}"
    for k in ga:
        assert same_bits(ga[k], gb[k]), f"geometry {k}"


@pytest.mark.parametrize("multi", [True, False], ids=["substeps", "one-step"])
def test_slab_step_equals_reference(tables2, oracle, ref, multi):
    tables = tables2 if multi else tables_multistep(tables2, False)
    tr, x, lo, hi, ex = slab_tracks(60_000, 5, len(SET2))
    a = oracle.electron_slab_step(tables, 77, tr, x, lo, hi, ex, nthreads=8)
    b = ref.electron_slab_step(tables, 77, tr, x, lo, hi, ex, nthreads=8)
    assert_same(a, b)
    nsub, left, res = a[2]["nsub"], a[2]["left"], a[1]
    if multi:
        # sub-steps do happen, and long chains of them
        assert (nsub > 1).mean() > 0.05 and nsub.max() >= 10
    else:
        assert (nsub == 1).all()
    assert 0.05 < left.mean() < 0.5
    # every branch of the loop's end is reached: MSC-limited, continuous,
    # ionisation, brem and annihilation winners, stops
    assert {-2, -1, 0, 1, 2} <= set(np.unique(res["winner"]).tolist())
    assert res["stopped"].sum() > 100


def test_slab_step_low_energy_and_thin_slabs(tables1, oracle, ref):
    """Tracks near the end of their range in thin slabs: stops, at-rest
    annihilation and short chains of sub-steps."""
    tr, x, lo, hi, ex = slab_tracks(40_000, 9, len(SET1), lnemin=math.log(1e-3), lnemax=math.log(0.3))
    hi = hi * 0.1
    x = x * 0.1
    a = oracle.electron_slab_step(tables1, 5, tr, x, lo, hi, ex, nthreads=8)
    b = ref.electron_slab_step(tables1, 5, tr, x, lo, hi, ex, nthreads=8)
    assert_same(a, b)
    assert a[1]["stopped"].mean() > 0.05
    pos = tr["ptype"] == 1
    assert (a[1]["nsec"][pos & (a[1]["stopped"] == 1)] == 2).all()   # e+ at rest: two gammas


def test_one_step_equals_howfar_geometry_perform(tables2, oracle):
    """Without sub-steps and with no way out of the world, the TrackElectron
    step is the manager's HowFar, the geometry cut and Perform
    (G4HepEmElectronManager.icc:35-45,461-483)."""
    tables = tables_multistep(tables2, False)
    tr, x, lo, hi, ex = slab_tracks(40_000, 21, len(SET2), exits=False)
    t1, r1, g1 = oracle.electron_slab_step(tables, 3, tr, x, lo, hi, ex, nthreads=8)
    # the same step through HowFar + geometry + Perform: the safety and the
    # distance to the plane along the direction
    tin = {k: np.array(v, copy=True) for k, v in tr.items()}
    onb = tin["on_boundary"].astype(bool)
    tin["safety"][:] = np.where(onb, 0.0, np.maximum(0.0, np.minimum(x - lo, hi - x)))
    ux = tin["dir"][: len(x)]
    with np.errstate(divide="ignore"):
        dist = np.where(ux > 0, (hi - x) / ux, np.where(ux < 0, (lo - x) / ux, np.inf))
    dist = dist + 0.0   # a track on the plane it moves out through: +0, as the slab step's move
    t2, r2 = oracle.step(tables, 3, tin, nthreads=8, geom_limit=dist)
    for k in OUT_FIELDS:
        if k != "displacement":   # applied to x by the slab step, reported by the manager's step
            assert same_bits(r1[k], r2[k]), k
    for k in TRACK_FIELDS:
        assert same_bits(t1[k], t2[k]), k


@pytest.mark.parametrize("e0,nev", [(1000.0, 100), (1.0e4, 24)])
def test_slab_shower_equals_reference(tables1, oracle, ref, e0, nev):
    """The oracle's slab shower against the shower stepped by the reference's
    own code, event by event, at 1 and 10 GeV."""
    from g4hepem_amd._native import slab_setup
    su = slab_setup(50, [2.3, 5.7], [0, 1], 0, e0, capacity=1 << 22, events_per_chunk=64)
    a = oracle.slab_shower(tables1, su, nev, 11, nthreads=8)
    b = ref.slab_shower(tables1, su, nev, 11, nthreads=8)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert (a[2], a[3]) == (b[2], b[3])
    bal = (a[0].sum(1) + a[1]) / 1e9
    assert np.abs(bal - e0).max() < 1e-4


def test_slab_shower_substeps_change_the_steps_not_the_energy(tables1, oracle):
    """The same showers with sub-stepping off: fewer sub-steps merge into
    fewer steps with it on, and both conserve energy."""
    from g4hepem_amd._native import slab_setup
    su = slab_setup(50, [2.3, 5.7], [0, 1], 0, 1000.0, capacity=1 << 22, events_per_chunk=64)
    on = oracle.slab_shower(tables1, su, 60, 4, nthreads=8)
    off = oracle.slab_shower(tables_multistep(tables1, False), su, 60, 4, nthreads=8)
    assert on[2] < off[2]
    for d, l, _, _ in (on, off):
        assert np.abs((d.sum(1) + l) / 1e9 - 1000.0).max() < 1e-4
