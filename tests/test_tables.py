"""Layout constants and internal consistency of the offline tables (the
reference pins the same constants in testing/G4HepEmDataInterfaces/
TestG4HepEmDataInterfaces.cc:39,84-86,119-121,236-239)."""
import numpy as np

from g4hepem_amd.data_layout import arr, params, view


def test_layout_constants(tables2):
    d = view(tables2)
    el = d.fTheElectronData.contents
    assert el.fELossEnergyGridSize == 85
    assert el.fENucEnergyGridSize == 128
    g = d.fTheGammaData.contents
    assert (g.fEGridSize0, g.fEGridSize1, g.fEGridSize2) == (32, 32, 256)
    assert g.fDataPerMat == 2 * 32 + 3 * 32 + 9 * 256 and g.fNumData0 == 64 and g.fNumData1 == 96
    sb = d.fTheSBTableData.contents
    assert (sb.fMaxZet, sb.fNumElEnergy, sb.fNumKappa) == (99, 65, 54)
    assert d.fTheElementData.contents.fMaxZet == 120
    p = params(tables2)
    assert p.fNumLossTableBins == 84 and p.fNumRegions == 1
    assert abs(p.fElectronTrackingCut - 1e-3) < 1e-15


def test_couples_materials_elements(tables2):
    d = view(tables2)
    mc = d.fTheMatCutData.contents
    assert mc.fNumMatCutData == 4
    zs = []
    for i in range(4):
        c = mc.fMatCutData[i]
        assert c.fHepEmMatIndex == i and c.fG4RegionIndex == 0
        assert c.fSecGamProdCutE > 0 and c.fSecElProdCutE >= 1e-3
        assert abs(c.fLogSecGamCutE - np.log(c.fSecGamProdCutE)) < 1e-15
        m = d.fTheMaterialData.contents.fMaterialData[i]
        zs.append(sorted(arr(m.fElementVect, m.fNumOfElement, np.int32).tolist()))
        assert m.fRadiationLength > 0 and m.fElectronDensity > 0
    assert zs == [[8, 74, 82], [82], [18], [14]]
    # Pb radiation length (PDG 5.612 mm) and lAr (140 mm) from the Tsai formula
    assert abs(d.fTheMaterialData.contents.fMaterialData[1].fRadiationLength - 5.612) < 0.05
    assert abs(d.fTheMaterialData.contents.fMaterialData[2].fRadiationLength - 140.0) < 2.0
    for z in (8, 14, 18, 74, 82):
        e = d.fTheElementData.contents.fElementData[z]
        assert e.fZet == z and abs(e.fZet13 ** 3 - z) < 1e-9


def test_eloss_tables_monotone(tables2):
    el = view(tables2).fTheElectronData.contents
    n = el.fELossEnergyGridSize
    data = arr(el.fELossData, 5 * n * el.fNumMatCuts)
    grid = arr(el.fELossEnergyGrid, n)
    assert np.all(np.diff(grid) > 0)
    for imc in range(el.fNumMatCuts):
        rng = data[5 * n * imc: 5 * n * imc + 2 * n: 2]
        dedx = data[5 * n * imc + 2 * n: 5 * n * imc + 4 * n: 2]
        assert np.all(np.diff(rng) > 0) and np.all(dedx > 0)


def test_xsec_blocks_and_selectors(tables2):
    el = view(tables2).fTheElectronData.contents
    xs = arr(el.fResMacXSecData, el.fResMacXSecNumData)
    for imc in range(el.fNumMatCuts):
        s = el.fResMacXSecStartIndexPerMatCut[imc]
        for _ in range(2):  # ioni then brem block
            m = int(xs[s])
            e = xs[s + 5: s + 5 + 3 * m: 3]
            v = xs[s + 6: s + 5 + 3 * m: 3]
            assert np.all(np.diff(e) > 0) and np.all(v >= 0)
            assert xs[s + 2] == v.max() and xs[s + 1] == e[np.argmax(v)]
            s += 5 + 3 * m
    # single-element couples have no selector; PbWO4 has 3-element CDFs in [0,1]
    assert el.fElemSelectorBremSBStartIndexPerMatCut[1] == -1
    st = el.fElemSelectorBremSBStartIndexPerMatCut[0]
    sel = arr(el.fElemSelectorBremSBData, el.fElemSelectorBremSBNumData)
    k, q = int(sel[st]), int(sel[st + 1])
    assert q == 3
    rows = sel[st + 4: st + 4 + k * q].reshape(k, q)
    assert np.all(np.diff(rows[:, 0]) > 0)
    assert np.all((rows[:, 1:] >= 0) & (rows[:, 1:] <= 1)) and np.all(np.diff(rows[:, 1:], axis=1) >= 0)


def test_sb_sampling_tables(tables2):
    sb = view(tables2).fTheSBTableData.contents
    data = arr(sb.fSBTableData, sb.fNumSBTableData)
    for z in (8, 14, 18, 74, 82):
        s = sb.fSBTablesStartPerZ[z]
        ndata, mn, mx, ncut = data[s: s + 4]
        assert ndata == (mx - mn + 1) * (ncut + 3 * 54) + 4
        size = int(ncut + 162)
        for ie in range(int(mx - mn + 1)):
            row = data[s + 4 + ie * size: s + 4 + (ie + 1) * size]
            cuts, cum = row[: int(ncut)], row[int(ncut)::3]
            assert cum[0] == 0.0 and cum[-1] == 1.0 and np.all(np.diff(cum) >= 0)
            assert np.all((cuts >= 0) & (cuts <= 1))


def test_digest_is_reproducible(tables2):
    from g4hepem_amd import Tables
    t = Tables(["G4_PbWO4", "G4_Pb", "G4_lAr", "G4_Si"], 0.7)
    assert t.digest() == tables2.digest()
    assert Tables(["G4_Pb"], 0.7).digest() != Tables(["G4_Pb"], 1.0).digest()
