// hepem_validate.cpp -- consistency check of a G4HepEmData (see
// hepem_validate.hh): every index that the stepping, the context packing or
// the device kernels derive from table contents is checked against the
// array that holds it.
#include "hepem_validate.hh"

#include <cmath>
#include <cstddef>

namespace {

// a[i] as an int in [lo, hi], for i inside an array of n values
bool IntAt(const double* a, long n, long i, long lo, long hi, int& out) {
  if (!a || i < 0 || i >= n) return false;
  const double v = a[i];
  if (!std::isfinite(v) || v < (double)lo || v > (double)hi || v != std::floor(v)) return false;
  out = (int)v;
  return true;
}

// one restricted cross-section block [n, maxE, maxV, logMin, invLD, (E, S, S'') x n]
// starting at s (G4HepEmElectronManager.icc:522-532); *end receives the index behind it
bool XsBlockOk(const double* xs, long size, long s, long& end) {
  int n = 0;
  if (!IntAt(xs, size, s, 2, 1 << 28, n)) return false;
  end = s + 5 + 3L * n;
  return end <= size;
}

// a brem element selector block [numData, numElem, logE0, invLD, numData x (E,
// cdf_0..cdf_{numElem-2})] of a material with nel elements
// (G4HepEmElectronInteractionBrem.icc:266-297)
bool BremSelectorOk(const double* d, long size, long s, int nel) {
  int numData = 0, numElem = 0;
  if (!IntAt(d, size, s, 2, 1 << 24, numData) || !IntAt(d, size, s + 1, 2, nel, numElem)) return false;
  return s + 4 + (long)numElem * numData <= size;
}

bool Sandia(int n, const double* e, const double* c) { return n >= 1 && e && c; }

}  // namespace

bool HepemTablesValid(const G4HepEmData* d) {
  if (!d || !d->fTheMatCutData || !d->fTheMaterialData || !d->fTheElementData || !d->fTheElectronData ||
      !d->fThePositronData || !d->fTheSBTableData || !d->fTheGammaData)
    return false;
  const G4HepEmMatCutData* mc = d->fTheMatCutData;
  const G4HepEmMaterialData* md = d->fTheMaterialData;
  const G4HepEmElementData* ed = d->fTheElementData;
  const int nmc = mc->fNumMatCutData, nmat = md->fNumMaterialData, maxZ = ed->fMaxZet;
  if (nmc < 1 || !mc->fMatCutData || nmat < 1 || !md->fMaterialData || maxZ < 1 || maxZ > 120 || !ed->fElementData)
    return false;
  for (int i = 0; i < nmc; ++i) {
    const int m = mc->fMatCutData[i].fHepEmMatIndex;
    if (m < 0 || m >= nmat) return false;
  }
  // materials: element lists with Z <= fMaxZet, photoabsorption intervals of
  // the material and of each of its elements
  for (int i = 0; i < nmat; ++i) {
    const G4HepEmMatData& m = md->fMaterialData[i];
    if (m.fNumOfElement < 1 || !m.fElementVect || !m.fNumOfAtomsPerVolumeVect) return false;
    if (!Sandia(m.fNumOfSandiaIntervals, m.fSandiaEnergies, m.fSandiaCoefficients)) return false;
    for (int k = 0; k < m.fNumOfElement; ++k) {
      const int z = m.fElementVect[k];
      if (z < 1 || z > maxZ) return false;
      const G4HepEmElemData& e = ed->fElementData[z];
      if (!Sandia(e.fNumOfSandiaIntervals, e.fSandiaEnergies, e.fSandiaCoefficients)) return false;
    }
  }
  // e-/e+: per-couple and per-material array sizes, cross-section blocks,
  // brem element selectors of the multi-element materials
  for (const G4HepEmElectronData* e : {d->fTheElectronData, d->fThePositronData}) {
    if (e->fNumMatCuts != nmc || e->fNumMaterials != nmat) return false;
    if (e->fELossEnergyGridSize < 2 || !e->fELossEnergyGrid || !e->fELossData) return false;
    if (e->fENucEnergyGridSize < 2 || !e->fENucEnergyGrid || !e->fENucMacXsecData || !e->fTr1MacXSecData) return false;
    if (!e->fResMacXSecStartIndexPerMatCut || !e->fResMacXSecData || e->fResMacXSecNumData < 1) return false;
    const long nxs = e->fResMacXSecNumData;
    for (int i = 0; i < nmc; ++i) {
      long end = 0;
      if (!XsBlockOk(e->fResMacXSecData, nxs, e->fResMacXSecStartIndexPerMatCut[i], end)) return false;  // ioni
      if (!XsBlockOk(e->fResMacXSecData, nxs, end, end)) return false;                                   // brem
      const int nel = md->fMaterialData[mc->fMatCutData[i].fHepEmMatIndex].fNumOfElement;
      if (nel < 2) continue;
      if (!e->fElemSelectorBremSBStartIndexPerMatCut || !e->fElemSelectorBremRBStartIndexPerMatCut) return false;
      if (!BremSelectorOk(e->fElemSelectorBremSBData, e->fElemSelectorBremSBNumData,
                          e->fElemSelectorBremSBStartIndexPerMatCut[i], nel) ||
          !BremSelectorOk(e->fElemSelectorBremRBData, e->fElemSelectorBremRBNumData,
                          e->fElemSelectorBremRBStartIndexPerMatCut[i], nel))
        return false;
    }
  }
  // Seltzer-Berger: the rows of every element in use and every couple's
  // gamma-cut indices into them (G4HepEmElectronInteractionBrem.icc:73-182)
  const G4HepEmSBTableData* sb = d->fTheSBTableData;
  if (sb->fNumKappa < 2 || sb->fNumElEnergy < 2 || sb->fNumElEnergy > 65 || !sb->fSBTableData ||
      sb->fNumHepEmMatCuts != nmc || !sb->fGammaCutIndxStartIndexPerMC || !sb->fGammaCutIndices)
    return false;
  const long nsb = sb->fNumSBTableData;
  for (int i = 0; i < nmc; ++i) {
    const G4HepEmMatData& m = md->fMaterialData[mc->fMatCutData[i].fHepEmMatIndex];
    const long c0 = sb->fGammaCutIndxStartIndexPerMC[i];
    if (c0 < 0 || c0 + m.fNumOfElement > sb->fNumElemsInMatCuts) return false;
    for (int k = 0; k < m.fNumOfElement; ++k) {
      const int z = m.fElementVect[k];
      const long s = sb->fSBTablesStartPerZ[z];
      int minE = 0, maxE = 0, ncut = 0;
      if (s < 0 || !IntAt(sb->fSBTableData, nsb, s + 1, 0, sb->fNumElEnergy - 1, minE) ||
          !IntAt(sb->fSBTableData, nsb, s + 2, minE, sb->fNumElEnergy - 1, maxE) ||
          !IntAt(sb->fSBTableData, nsb, s + 3, 1, 1 << 20, ncut))
        return false;
      if (s + 4 + (long)(maxE - minE + 1) * (ncut + 3L * sb->fNumKappa) > nsb) return false;
      const int ic = sb->fGammaCutIndices[c0 + k];
      if (ic < 0 || ic >= ncut) return false;
    }
  }
  // gamma: block sizes per material and the conversion element selector
  // (G4HepEmGammaManager.icc:108-153, G4HepEmGammaInteractionConversion.icc:151-177)
  const G4HepEmGammaData* g = d->fTheGammaData;
  if (g->fNumMaterials != nmat || !g->fMacXsecData || g->fEGridSize0 < 2 || g->fEGridSize1 < 2 || g->fEGridSize2 < 2 ||
      g->fNumData0 != 2 * g->fEGridSize0 || g->fNumData1 != 3 * g->fEGridSize1 ||
      g->fDataPerMat != g->fNumData0 + g->fNumData1 + 9 * g->fEGridSize2 || !g->fElemSelectorConvStartIndexPerMat)
    return false;
  for (int i = 0; i < nmat; ++i) {
    const int nel = md->fMaterialData[i].fNumOfElement;
    if (nel < 2) continue;
    const long s = g->fElemSelectorConvStartIndexPerMat[i];
    int numElem = 0;
    if (g->fElemSelectorConvEgridSize < 2 || !g->fElemSelectorConvEgrid ||
        !IntAt(g->fElemSelectorConvData, g->fElemSelectorConvNumData, s, 2, nel, numElem))
      return false;
    if (s + 1 + (long)g->fElemSelectorConvEgridSize * (numElem - 1) > g->fElemSelectorConvNumData) return false;
  }
  return true;
}
