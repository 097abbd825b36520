"""ctypes mirrors of include/g4hepem_amd_data.h (the reference-layout table
structs), for reading tables from Python."""
from __future__ import annotations

import ctypes as C

import numpy as np

d, i, b, vp = C.c_double, C.c_int, C.c_bool, C.c_void_p
pd, pi = C.POINTER(C.c_double), C.POINTER(C.c_int)


class RegionParameters(C.Structure):
    _fields_ = [("fFinalRange", d), ("fDRoverRange", d), ("fLinELossLimit", d), ("fMSCRangeFactor", d),
                ("fMSCSafetyFactor", d), ("fIsMSCMinimalStepLimit", b), ("fIsELossFluctuation", b),
                ("fIsMultipleStepsInMSCTrans", b), ("fIsApplyCuts", b)]


class Parameters(C.Structure):
    _fields_ = [("fElectronTrackingCut", d), ("fGammaTrackingCut", d), ("fMinLossTableEnergy", d),
                ("fMaxLossTableEnergy", d), ("fNumLossTableBins", i), ("fElectronBremModelLim", d),
                ("fIsMSCPositronCor", b), ("fIsMSCDisplacement", b), ("fNumRegions", i),
                ("fParametersPerRegion", C.POINTER(RegionParameters))]


class MCCData(C.Structure):
    _fields_ = [("fSecElProdCutE", d), ("fSecPosProdCutE", d), ("fSecGamProdCutE", d), ("fLogSecGamCutE", d),
                ("fHepEmMatIndex", i), ("fG4MatCutIndex", i), ("fG4RegionIndex", i)]


class MatCutData(C.Structure):
    _fields_ = [("fNumG4MatCuts", i), ("fNumMatCutData", i), ("fG4MCIndexToHepEmMCIndex", pi),
                ("fMatCutData", C.POINTER(MCCData))]


class MatData(C.Structure):
    _fields_ = [("fG4MatIndex", i), ("fNumOfElement", i), ("fElementVect", pi), ("fNumOfAtomsPerVolumeVect", pd),
                ("fDensity", d), ("fDensityCorFactor", d), ("fElectronDensity", d), ("fRadiationLength", d),
                ("fMeanExEnergy", d), ("fNumOfSandiaIntervals", i), ("fSandiaEnergies", pd),
                ("fSandiaCoefficients", pd), ("fZeff", d), ("fZeff23", d), ("fZeffSqrt", d), ("fUMSCPar", d),
                ("fUMSCStepMinPars", d * 2), ("fUMSCTailCoeff", d * 4), ("fUMSCThetaCoeff", d * 2)]


class MaterialData(C.Structure):
    _fields_ = [("fNumG4Material", i), ("fNumMaterialData", i), ("fG4MatIndexToHepEmMatIndex", pi),
                ("fMaterialData", C.POINTER(MatData))]


class ElemData(C.Structure):
    _fields_ = [("fZet", d), ("fZet13", d), ("fZet23", d), ("fCoulomb", d), ("fLogZ", d), ("fZFactor1", d),
                ("fDeltaMaxLow", d), ("fDeltaMaxHigh", d), ("fILVarS1", d), ("fILVarS1Cond", d),
                ("fNumOfSandiaIntervals", i), ("fSandiaEnergies", pd), ("fSandiaCoefficients", pd),
                ("fKShellBindingEnergy", d)]


class ElementData(C.Structure):
    _fields_ = [("fMaxZet", i), ("fElementData", C.POINTER(ElemData))]


class ElectronData(C.Structure):
    _fields_ = [("fNumMatCuts", i), ("fNumMaterials", i), ("fELossEnergyGridSize", i), ("fELossLogMinEkin", d),
                ("fELossEILDelta", d), ("fELossEnergyGrid", pd), ("fELossData", pd), ("fResMacXSecNumData", i),
                ("fResMacXSecStartIndexPerMatCut", pi), ("fResMacXSecData", pd), ("fENucEnergyGridSize", i),
                ("fENucLogMinEkin", d), ("fENucEILDelta", d), ("fENucEnergyGrid", pd), ("fENucMacXsecData", pd),
                ("fTr1MacXSecData", pd), ("fElemSelectorIoniNumData", i), ("fElemSelectorIoniStartIndexPerMatCut", pi),
                ("fElemSelectorIoniData", pd), ("fElemSelectorBremSBNumData", i),
                ("fElemSelectorBremSBStartIndexPerMatCut", pi), ("fElemSelectorBremSBData", pd),
                ("fElemSelectorBremRBNumData", i), ("fElemSelectorBremRBStartIndexPerMatCut", pi),
                ("fElemSelectorBremRBData", pd)]


class SBTableData(C.Structure):
    _fields_ = [("fMaxZet", i), ("fNumElEnergy", i), ("fNumKappa", i), ("fLogMinElEnergy", d),
                ("fILDeltaElEnergy", d), ("fElEnergyVect", d * 65), ("fLElEnergyVect", d * 65),
                ("fKappaVect", d * 54), ("fLKappaVect", d * 54), ("fNumHepEmMatCuts", i), ("fNumElemsInMatCuts", i),
                ("fGammaCutIndxStartIndexPerMC", pi), ("fGammaCutIndices", pi), ("fNumSBTableData", i),
                ("fSBTablesStartPerZ", i * 121), ("fSBTableData", pd)]


class GammaData(C.Structure):
    _fields_ = [("fNumMaterials", i), ("fEGridSize0", i), ("fEGridSize1", i), ("fEGridSize2", i), ("fDataPerMat", i),
                ("fNumData0", i), ("fNumData1", i), ("fEMin0", d), ("fEMax0", d), ("fLogEMin0", d), ("fEILDelta0", d),
                ("fEMax1", d), ("fLogEMin1", d), ("fEILDelta1", d), ("fEMax2", d), ("fLogEMin2", d), ("fEILDelta2", d),
                ("fMacXsecData", pd), ("fElemSelectorConvEgridSize", i), ("fElemSelectorConvNumData", i),
                ("fElemSelectorConvLogMinEkin", d), ("fElemSelectorConvEILDelta", d),
                ("fElemSelectorConvStartIndexPerMat", pi), ("fElemSelectorConvEgrid", pd),
                ("fElemSelectorConvData", pd)]


class HepEmData(C.Structure):
    _fields_ = [("fTheMatCutData", C.POINTER(MatCutData)), ("fTheMaterialData", C.POINTER(MaterialData)),
                ("fTheElementData", C.POINTER(ElementData)), ("fTheElectronData", C.POINTER(ElectronData)),
                ("fThePositronData", C.POINTER(ElectronData)), ("fTheSBTableData", C.POINTER(SBTableData)),
                ("fTheGammaData", C.POINTER(GammaData))]


def view(tables) -> HepEmData:
    return C.cast(tables.data, C.POINTER(HepEmData)).contents


def params(tables) -> Parameters:
    return C.cast(tables.pars, C.POINTER(Parameters)).contents


def arr(ptr, n, dtype=np.float64) -> np.ndarray:
    """Copy of n values behind a ctypes pointer."""
    return np.ctypeslib.as_array(ptr, (int(n),)).astype(dtype, copy=True)
