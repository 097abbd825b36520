"""The oracle (C restatement of the reference run-time physics) against the
golden vectors written by the compiled reference, against the reference itself
where it is built here, and against what mathematics fixes."""
import math

import numpy as np
import pytest

from conftest import STEP_SEED, golden_tracks, same_bits
from oracle.oracle_py import OUT_FIELDS, TRACK_FIELDS, ref_available


# ---- golden vectors ------------------------------------------------------------------------

def test_tables_match_golden_digest(golden, tables1, tables2):
    assert tables1.digest() == str(golden["digest_set1"])
    assert tables2.digest() == str(golden["digest_set2"])


@pytest.mark.parametrize("pt,name", [(0, "e"), (1, "p")])
def test_lookup_golden(golden, tables1, oracle, pt, name):
    out = oracle.lookup(tables1, pt, golden[f"c1_{name}_imc"], golden[f"c1_{name}_ekin"], golden[f"c1_{name}_lekin"])
    assert same_bits(out, golden[f"c1_{name}_out"])


@pytest.mark.parametrize("pt,name", [(0, "e"), (1, "p")])
def test_electron_probes_golden(golden, tables2, oracle, pt, name):
    imc, e, le = golden[f"pe_{name}_imc"], golden[f"pe_{name}_ekin"], golden[f"pe_{name}_lekin"]
    for w in range(6):
        assert same_bits(oracle.probe_electron(tables2, pt, w, imc, e, le), golden[f"pe_{name}_out"][w]), w


@pytest.mark.parametrize("which", [0, 1, 2])
def test_selectors_golden(golden, tables2, oracle, which):
    le, u = golden[f"sel{which}_lekin"], golden[f"sel{which}_u"]
    imc = np.zeros(len(le), dtype=np.int32)
    out = oracle.probe_selector(tables2, 0, which, imc, np.exp(le), le, u)
    assert same_bits(out, golden[f"sel{which}_out"])
    assert set(np.unique(out)) <= {0, 1, 2}


def test_gamma_golden(golden, tables2, oracle):
    xs, pe, pid = oracle.probe_gamma(tables2, golden["gm_imc"], golden["gm_ekin"], golden["gm_lekin"], golden["gm_u"])
    assert same_bits(xs, golden["gm_xs"])
    assert same_bits(pid, golden["gm_pid"])
    known = pid != 3  # the reference's PE value is not defined in the gamma-nuclear bin
    assert same_bits(pe[known], golden["gm_pe"][known])
    # all three energy windows are exercised
    e = golden["gm_ekin"]
    assert (e < 0.15).sum() > 100 and ((e > 0.15) & (e < 1.022)).sum() > 30 and (e > 1.022).sum() > 100


@pytest.mark.parametrize("r", [0, 1, 2])
def test_step_golden(golden, tables2, oracle, r):
    tin = golden_tracks(golden, f"st{r}_in")
    tout, res = oracle.step(tables2, STEP_SEED, tin, nthreads=4)
    for k in OUT_FIELDS:
        assert same_bits(res[k], golden[f"st{r}_res_{k}"]), k
    el = tin["ptype"] != 2
    for k in TRACK_FIELDS:
        a, b = tout[k], golden[f"st{r}_out_{k}"]
        if k.startswith("msc"):
            a, b = a[el], b[el]
        assert same_bits(a, b), k


def test_step_chain_golden(golden, tables2, oracle):
    cur = golden_tracks(golden, "ch_in")
    for s in range(4):
        alive = cur["ekin"] > 0
        cur, res = oracle.step(tables2, STEP_SEED + 1, cur, nthreads=2)
        for k, (dt, c) in OUT_FIELDS.items():
            m = np.tile(alive, c)
            assert same_bits(res[k][m], golden[f"ch{s}_res_{k}"][m]), (s, k)
        for k in ("ekin", "dir", "num_ia_left", "rng_ctr", "rng_gauss", "rng_has_gauss"):
            m = np.tile(alive, TRACK_FIELDS[k][1])
            assert same_bits(cur[k][m], golden[f"ch{s}_out_{k}"][m]), (s, k)
    # the chain exercises boundaries, cached Gaussians and non-first MSC steps
    assert golden["ch_in_on_boundary"].sum() > 50
    assert golden["ch1_out_rng_has_gauss"].sum() > 0


def test_stream_high_word_golden(golden, tables2, oracle):
    """Tracks selecting their Philox stream by the 64-bit pair (rng_stream, id)."""
    tin = golden_tracks(golden, "hs_in")
    assert np.count_nonzero(tin["rng_stream"]) > 500
    tout, res = oracle.step(tables2, STEP_SEED + 3, tin, nthreads=4)
    for k in OUT_FIELDS:
        assert same_bits(res[k], golden[f"hs_res_{k}"]), k
    for k in ("ekin", "dir", "num_ia_left", "rng_ctr", "rng_gauss", "rng_stream"):
        assert same_bits(tout[k], golden[f"hs_out_{k}"]), k
    # the high word changes the stream: the same tracks with rng_stream = 0 differ
    t0 = dict(tin)
    t0["rng_stream"] = np.zeros_like(tin["rng_stream"])
    _, res0 = oracle.step(tables2, STEP_SEED + 3, t0, nthreads=4)
    assert not same_bits(res0["edep"], res["edep"])


GEOM_STATE = ("ekin", "dir", "num_ia_left", "rng_ctr", "rng_gauss", "rng_has_gauss", "msc_init_range",
              "msc_dyn_range_factor", "msc_tlimit_min", "msc_first_step", "on_boundary")


def test_geometry_chain_golden(golden_geom, tables2, oracle):
    """Steps cut by a geometry between HowFar and Perform, chained over 3 steps."""
    assert tables2.digest() == str(golden_geom["digest_set2"])
    cur = golden_tracks(golden_geom, "gc_in")
    for s in range(3):
        alive = cur["ekin"] > 0
        lim = golden_geom[f"gc{s}_limit"]
        cur, res = oracle.step(tables2, STEP_SEED + 2, cur, nthreads=4, geom_limit=lim)
        for k, (dt, c) in OUT_FIELDS.items():
            m = np.tile(alive, c)
            assert same_bits(res[k][m], golden_geom[f"gc{s}_res_{k}"][m]), (s, k)
        for k in GEOM_STATE:
            m = np.tile(alive, TRACK_FIELDS[k][1])
            if k.startswith("msc"):
                m &= np.tile(cur["ptype"] != 2, TRACK_FIELDS[k][1])
            assert same_bits(cur[k][m], golden_geom[f"gc{s}_out_{k}"][m]), (s, k)
        # a cut step ends on the boundary at exactly the cut distance
        onb = cur["on_boundary"][alive] == 1
        assert 0.1 < onb.mean() < 0.9, onb.mean()
        assert np.all(res["geom_step"][alive][onb] == lim[alive][onb])
        assert np.all(res["geom_step"][alive] <= lim[alive])
        # no discrete interaction and no secondaries on a boundary step
        assert np.all(res["nsec"][alive][onb & (cur["ptype"][alive] == 2)] == 0)


@pytest.mark.skipif(not ref_available(), reason="reference tree not built here")
def test_oracle_equals_reference_live_geometry(tables2, oracle):
    from g4hepem_amd.tracks import step_tracks
    from oracle.oracle_py import Ref
    ref = Ref()
    tr = step_tracks(20000, 778, 4, lnemin=math.log(1e-3), lnemax=math.log(1e6))
    lim = np.exp(np.random.Generator(np.random.Philox(779)).uniform(math.log(1e-4), math.log(20.0), 20000))
    a_t, a_r = oracle.step(tables2, 4243, tr, 8, geom_limit=lim)
    b_t, b_r = ref.step(tables2, 4243, tr, 8, geom_limit=lim)
    for k in OUT_FIELDS:
        assert same_bits(a_r[k], b_r[k]), k
    for k in ("ekin", "dir", "num_ia_left", "rng_ctr", "rng_gauss", "on_boundary"):
        assert same_bits(a_t[k], b_t[k]), k


@pytest.mark.skipif(not ref_available(), reason="reference tree not built here")
def test_oracle_equals_reference_live(tables2, oracle):
    from g4hepem_amd.tracks import step_tracks
    from oracle.oracle_py import Ref
    ref = Ref()
    tr = step_tracks(20000, 777, 4, lnemin=math.log(1e-3), lnemax=math.log(1e8))
    a_t, a_r = oracle.step(tables2, 4242, tr, 8)
    b_t, b_r = ref.step(tables2, 4242, tr, 8)
    for k in OUT_FIELDS:
        assert same_bits(a_r[k], b_r[k]), k
    for k in ("ekin", "dir", "num_ia_left", "rng_ctr", "rng_gauss"):
        assert same_bits(a_t[k], b_t[k]), k


# ---- pins from mathematics -------------------------------------------------------------------

def test_philox_known_answers(oracle):
    # Random123 known-answer vectors for philox4x32-10
    assert list(oracle.philox([0, 0, 0, 0], [0, 0])) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert list(oracle.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2)) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert list(oracle.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0])) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_uniform_stream(oracle):
    u = oracle.uniforms(0x1234, 7, 0, 200000)
    assert u.min() > 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.003
    assert abs(u.var() - 1.0 / 12.0) < 0.001
    # draws are a pure function of (seed, id, index): any window reproduces
    assert same_bits(oracle.uniforms(0x1234, 7, 1001, 10), u[1001:1011])
    assert not np.array_equal(oracle.uniforms(0x1234, 8, 0, 10), u[:10])


def test_vdt_exp_log_accuracy(oracle):
    xs = np.concatenate([np.linspace(-700, 700, 2001), np.linspace(-1, 1, 1001)])
    for x in xs:
        e = oracle.vdt_exp(x)
        assert abs(e - math.exp(x)) <= 3e-16 * math.exp(x), x
    for x in np.geomspace(1e-300, 1e300, 3001):
        l = oracle.vdt_log(x)
        assert abs(l - math.log(x)) <= 4e-16 * max(1.0, abs(math.log(x))), x
    assert oracle.vdt_exp(0.0) == 1.0 and oracle.vdt_log(1.0) == 0.0
    assert oracle.vdt_exp(800.0) == math.inf and oracle.vdt_exp(-800.0) == 0.0
    assert math.isnan(oracle.vdt_log(-1.0))


def test_spline_exact_at_nodes_and_inverse_range(tables1, oracle):
    """At grid nodes the spline returns the node values (b = 0), and
    GetInvRange(GetRestRange(E)) reproduces E (the composition the reference's
    own ELossData test evaluates)."""
    from g4hepem_amd.data_layout import arr, view
    ed = view(tables1).fTheElectronData.contents
    n = ed.fELossEnergyGridSize
    assert n == 85
    grid = arr(ed.fELossEnergyGrid, n)
    data = arr(ed.fELossData, 5 * n * ed.fNumMatCuts)
    for imc in (0, 1):
        imcs = np.full(n - 1, imc, dtype=np.int32)
        e = grid[:-1]
        out = oracle.lookup(tables1, 0, imcs, e, np.log(e))
        rng = data[5 * n * imc: 5 * n * imc + 2 * n: 2][:-1]
        dedx = data[5 * n * imc + 2 * n: 5 * n * imc + 4 * n: 2][:-1]
        assert np.allclose(out[1], rng, rtol=1e-13, atol=0)
        assert np.allclose(out[0], dedx, rtol=1e-13, atol=0)
        # inverse range at the range nodes returns the energy nodes
        assert np.allclose(out[2], e, rtol=1e-12)
    from g4hepem_amd.tracks import lookup_queries
    imc, e, le = lookup_queries(5000, 9, 2, lnemin=math.log(1e-3), lnemax=math.log(1e7))
    out = oracle.lookup(tables1, 0, imc, e, le)
    assert np.max(np.abs(out[2] / e - 1.0)) < 2e-3


def test_annihilation_closed_form(tables2, oracle):
    """Heitler e+e- -> 2 gamma cross section against an independent evaluation."""
    from g4hepem_amd.data_layout import view
    e = np.geomspace(1e-3, 1e5, 200)
    imc = np.full(200, 1, dtype=np.int32)  # G4_Pb
    xs = oracle.probe_electron(tables2, 1, 4, imc, e, np.log(e))
    ne = view(tables2).fTheMaterialData.contents.fMaterialData[1].fElectronDensity
    pir02 = math.pi * 2.8179403e-12 ** 2
    for k in range(200):
        g = e[k] / 0.51099891 + 1.0
        ref = ne * pir02 / (g + 1.0) * ((g * g + 4 * g + 1) / (g * g - 1) * math.log(g + math.sqrt(g * g - 1))
                                         - (g + 3) / math.sqrt(g * g - 1))
        assert abs(xs[k] / ref - 1.0) < 1e-5, e[k]
    xs8 = oracle.probe_electron(tables2, 1, 5, imc, e, np.log(e))
    xsr = oracle.probe_electron(tables2, 1, 4, imc, 0.8 * e, np.log(0.8 * e))
    assert np.allclose(xs8, xsr, rtol=1e-14)


def test_step_conservation_and_invariants(tables2, oracle):
    from g4hepem_amd.tracks import step_tracks
    n = 30000
    tr = step_tracks(n, 31, 4, lnemin=math.log(1e-3), lnemax=math.log(1e5))
    out, res = oracle.step(tables2, 77, tr, 8)
    check_step_invariants(tr, out, res)


def check_step_invariants(tr, out, res, tol=1e-9, dir_tol=1e-12):
    """Energy conservation per track, unit directions, process ids, RNG use."""
    n = len(tr["ekin"])
    me = 0.51099891
    pt = tr["ptype"]
    e_in = tr["ekin"]
    valid = np.stack([res["nsec"] > 0, res["nsec"] > 1])
    secE = np.where(valid, res["sec_ekin"].reshape(2, n), 0.0).sum(0)
    sec_pt = np.where(valid, res["sec_ptype"].reshape(2, n), -1)
    # e+ annihilation creates 2 m c^2 of photon energy
    ann = (pt == 1) & (res["nsec"] == 2) & (sec_pt[0] == 2) & (sec_pt[1] == 2)
    # pair production turns 2 m c^2 of photon energy into rest mass
    pair = (pt == 2) & (res["winner"] == 0) & (res["nsec"] == 2)
    lhs = e_in + np.where(ann, 2 * me, 0.0)
    rhs = out["ekin"] + res["edep"] + secE + np.where(pair, 2 * me, 0.0)
    assert np.all(np.abs(lhs - rhs) <= tol * np.maximum(lhs, 1e-3)), np.max(np.abs(lhs - rhs) / lhs)
    d = out["dir"].reshape(3, n)
    alive = out["ekin"] > 0
    assert np.all(np.abs((d[:, alive] ** 2).sum(0) - 1.0) < dir_tol)
    sd = res["sec_dir"].reshape(3, 2, n)
    for s in range(2):
        has = res["nsec"] > s
        assert np.all(np.abs((sd[:, s, has] ** 2).sum(0) - 1.0) < dir_tol)
    w = res["winner"]
    assert np.all(np.isin(w[pt != 2], [-2, -1, 0, 1, 2, 3]))
    assert np.all(np.isin(w[pt == 2], [0, 1, 2, 3]))
    assert np.all(w[pt == 0] != 2)  # no annihilation for e-
    assert np.all(out["rng_ctr"] > tr["rng_ctr"])
    assert np.all(res["nsec"] <= 2)
    assert np.all(res["true_step"] >= res["geom_step"] * (1 - 1e-15))
    # Moller: secondary takes at most half of the pre-interaction energy
    moller = (pt == 0) & (w == 0) & (res["nsec"] == 1)
    assert np.all(res["sec_ekin"][:n][moller] <= out["ekin"][moller] * (1 + 1e-12))
