"""The nuclear branches of the stepping (e-/e+ winner 3, gamma process 3) on
tables with non-zero lepto- and photo-nuclear cross sections
(Tables.set_nuclear): the oracle against golden vectors that the compiled
reference wrote on the same tables (tests/golden/golden_nuc_v1.npz), against
the compiled reference live, and against what the stored tables fix.

Reference code on this path: G4HepEmElectronManager.icc:85 (nuclear MFP in
HowFarToDiscreteInteraction), :418-421 (CheckDelta), :534-541,570-574
(GetMacXSecNuclear[ForStepping]); G4HepEmGammaManager.icc:199-214 (the
gamma-nuclear bin of SampleInteraction)."""
import math

import numpy as np
import pytest

from conftest import CUT, SET2, STEP_SEED, golden_tracks, same_bits
from oracle.oracle_py import OUT_FIELDS, TRACK_FIELDS, ref_available

GOLDEN_NUC = __import__("pathlib").Path(__file__).resolve().parent / "golden" / "golden_nuc_v1.npz"
NUC_SCALES = (0.5, 0.2)


@pytest.fixture(scope="module")
def gnuc():
    return np.load(GOLDEN_NUC, allow_pickle=False)


@pytest.fixture(scope="module")
def tnuc():
    from g4hepem_amd import Tables
    return Tables(SET2, CUT).set_nuclear(*NUC_SCALES)


def enuc_shape(e):
    return np.where(e < 100.0, 0.0, 1.0 + 0.25 * np.log(e / 100.0))


def gnuc_shape(e):
    t = (e - 5.0) ** 2
    return np.where(e <= 5.0, 0.0, t / (t + 4.0) * (0.2 + 25.0 / ((e - 13.0) ** 2 + 25.0)))


def test_nuclear_tables_digest(gnuc, tnuc, tables2):
    assert tnuc.digest() == str(gnuc["digest_nuc"])
    assert tnuc.digest() != tables2.digest()


def test_zero_scales_restore_the_built_tables(tables2):
    from g4hepem_amd import Tables
    t = Tables(SET2, CUT).set_nuclear(0.3, 0.4).set_nuclear(0.0, 0.0)
    assert t.digest() == tables2.digest()


def test_set_nuclear_rejects_bad_scales():
    from g4hepem_amd import HepEmError, Tables
    t = Tables(SET2, CUT)
    for bad in ((-1.0, 0.0), (0.0, float("nan")), (float("inf"), 0.0)):
        with pytest.raises(HepEmError):
            t.set_nuclear(*bad)


@pytest.mark.parametrize("pt,name", [(0, "e"), (1, "p")])
def test_electron_probes_nuclear_golden(gnuc, tnuc, oracle, pt, name):
    imc, e, le = gnuc[f"pe_{name}_imc"], gnuc[f"pe_{name}_ekin"], gnuc[f"pe_{name}_lekin"]
    for w in range(6):
        assert same_bits(oracle.probe_electron(tnuc, pt, w, imc, e, le), gnuc[f"pe_{name}_out"][w]), w
    nuc = gnuc[f"pe_{name}_out"][3]
    assert (nuc > 0).mean() > 0.5
    assert np.all(nuc[e < 100.0] == 0.0)


def test_gamma_nuclear_golden(gnuc, tnuc, oracle):
    xs, pe, pid = oracle.probe_gamma(tnuc, gnuc["gm_imc"], gnuc["gm_ekin"], gnuc["gm_lekin"], gnuc["gm_u"])
    assert same_bits(xs, gnuc["gm_xs"])
    assert same_bits(pid, gnuc["gm_pid"])
    known = pid != 3  # the reference's PE value is not defined in the gamma-nuclear bin
    assert same_bits(pe[known], gnuc["gm_pe"][known])
    assert (pid == 3).mean() > 0.01
    # no gamma-nuclear below the 5 MeV threshold of the table shape
    assert np.all(pid[gnuc["gm_ekin"] < 5.0] != 3)


@pytest.mark.parametrize("r", [0, 1])
def test_step_nuclear_golden(gnuc, tnuc, oracle, r):
    tin = golden_tracks(gnuc, f"st{r}_in")
    tout, res = oracle.step(tnuc, STEP_SEED + 4, tin, nthreads=4)
    for k in OUT_FIELDS:
        assert same_bits(res[k], gnuc[f"st{r}_res_{k}"]), k
    el = tin["ptype"] != 2
    for k in TRACK_FIELDS:
        a, b = tout[k], gnuc[f"st{r}_out_{k}"]
        if k.startswith("msc"):
            a, b = a[el], b[el]
        assert same_bits(a, b), k
    w = res["winner"]
    # both nuclear winners occur in at least 1 % of the tracks of their kind
    assert ((w == 3) & el).sum() >= 0.01 * el.sum()
    assert ((w == 3) & ~el).sum() >= 0.01 * (~el).sum()
    # a nuclear step creates no secondaries and resets that interaction length
    n = len(w)
    nuc = w == 3
    assert np.all(res["nsec"][nuc] == 0)
    nial = tout["num_ia_left"].reshape(4, n)
    assert np.all(nial[3][nuc & el] == -1.0)
    assert np.all(nial[0][nuc & ~el] == -1.0)


def test_step_chain_nuclear_golden(gnuc, tnuc, oracle):
    cur = golden_tracks(gnuc, "ch_in")
    seen = 0
    for s in range(3):
        alive = cur["ekin"] > 0
        cur, res = oracle.step(tnuc, STEP_SEED + 5, cur, nthreads=2)
        for k, (dt, c) in OUT_FIELDS.items():
            m = np.tile(alive, c)
            assert same_bits(res[k][m], gnuc[f"ch{s}_res_{k}"][m]), (s, k)
        for k in ("ekin", "dir", "num_ia_left", "rng_ctr", "rng_gauss", "rng_has_gauss"):
            m = np.tile(alive, TRACK_FIELDS[k][1])
            assert same_bits(cur[k][m], gnuc[f"ch{s}_out_{k}"][m]), (s, k)
        seen += int((res["winner"][alive] == 3).sum())
    assert seen > 20


@pytest.mark.skipif(not ref_available(), reason="reference tree not built here")
def test_oracle_equals_reference_live_nuclear(tnuc, oracle):
    from g4hepem_amd.tracks import step_tracks
    from oracle.oracle_py import Ref
    ref = Ref()
    tr = step_tracks(20000, 877, 4, lnemin=math.log(1e-1), lnemax=math.log(1e6))
    a_t, a_r = oracle.step(tnuc, 4244, tr, 8)
    b_t, b_r = ref.step(tnuc, 4244, tr, 8)
    for k in OUT_FIELDS:
        assert same_bits(a_r[k], b_r[k]), k
    for k in ("ekin", "dir", "num_ia_left", "rng_ctr", "rng_gauss"):
        assert same_bits(a_t[k], b_t[k]), k
    assert (a_r["winner"] == 3).sum() > 300


# ---- pins from the stored tables ---------------------------------------------------------------

def test_nuclear_xsec_at_grid_nodes_per_material(tnuc, oracle):
    """At its grid nodes GetMacXSecNuclear returns the stored node values,
    which are enuc_scale/X0 times the energy shape of the material the couple
    points at: a wrong material offset, a dropped clamp or the wrong particle's
    table fails here."""
    from g4hepem_amd.data_layout import arr, view
    d = view(tnuc)
    mcc = d.fTheMatCutData.contents
    for pt, ed in ((0, d.fTheElectronData.contents), (1, d.fThePositronData.contents)):
        n = ed.fENucEnergyGridSize
        grid = arr(ed.fENucEnergyGrid, n)
        for imc in range(len(SET2)):
            imat = mcc.fMatCutData[imc].fHepEmMatIndex
            x0 = d.fTheMaterialData.contents.fMaterialData[imat].fRadiationLength
            e = grid[:-1]
            got = oracle.probe_electron(tnuc, pt, 3, np.full(n - 1, imc, dtype=np.int32), e, np.log(e))
            want = NUC_SCALES[0] / x0 * enuc_shape(e)
            assert np.allclose(got, want, rtol=1e-12, atol=0), (pt, imc)
            # below the grid: zero
            lo = np.array([50.0, 99.0])
            assert np.all(oracle.probe_electron(tnuc, pt, 3, np.full(2, imc, dtype=np.int32), lo, np.log(lo)) == 0.0)


def test_gamma_nuclear_fraction_matches_tables(tnuc, oracle):
    """At a window-2 node the total is conversion + Compton + PE + the
    gamma-nuclear part, and SampleInteraction returns 3 exactly for
    u > (conv + Compton + PE)/total: the frequency of process 3 over a
    uniform grid of u fixes the last bin of the cumulative scan."""
    from g4hepem_amd.data_layout import arr, view
    d = view(tnuc)
    gd = d.fTheGammaData.contents
    n2 = gd.fEGridSize2
    for imc in (0, 1, 2):
        imat = d.fTheMatCutData.contents.fMatCutData[imc].fHepEmMatIndex
        x0 = d.fTheMaterialData.contents.fMaterialData[imat].fRadiationLength
        blk = arr(gd.fMacXsecData, gd.fNumMaterials * gd.fDataPerMat)[imat * gd.fDataPerMat + gd.fNumData0 + gd.fNumData1:]
        rows = blk[: 9 * n2].reshape(n2, 9)
        for node in (20, 40, 120):
            e = rows[node, 0]
            tot, conv, comp, pe = rows[node, 1], rows[node, 3], rows[node, 5], rows[node, 7]
            assert abs(tot - (conv + comp + pe + NUC_SCALES[1] / x0 * gnuc_shape(e))) <= 1e-14 * tot
            m = 4000
            u = (np.arange(m) + 0.5) / m
            xs, _, pid = oracle.probe_gamma(tnuc, np.full(m, imc, dtype=np.int32), np.full(m, e), np.full(m, math.log(e)), u)
            assert np.allclose(xs, tot, rtol=1e-13)
            frac3 = (pid == 3).mean()
            want = 1.0 - (conv + comp + pe) / tot
            assert abs(frac3 - want) <= 1.0 / m + 1e-12, (imc, node, frac3, want)
            # the three other processes fill the first bins in order
            assert np.all(np.diff(pid) >= 0)
