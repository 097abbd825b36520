/*
 * g4hepem_amd_data.h -- plain-C declarations of the G4HepEm table structures.
 *
 * These structs are byte-for-byte layout-identical to the reference's
 * (non-CUDA build) data structures, so that a caller holding a reference
 * G4HepEmData / G4HepEmParameters can hand it to this library unchanged:
 *
 *   G4HepEmData               G4HepEm/G4HepEmData/include/G4HepEmData.hh:44-74
 *   G4HepEmParameters         G4HepEm/G4HepEmData/include/G4HepEmParameters.hh:52-93
 *   G4HepEmRegionParmeters    G4HepEm/G4HepEmData/include/G4HepEmParameters.hh:25-49
 *   G4HepEmMatCutData/MCCData G4HepEm/G4HepEmData/include/G4HepEmMatCutData.hh:41-68
 *   G4HepEmMaterialData/MatData G4HepEm/G4HepEmData/include/G4HepEmMaterialData.hh:41-95
 *   G4HepEmElementData/ElemData G4HepEm/G4HepEmData/include/G4HepEmElementData.hh:36-86
 *   G4HepEmElectronData       G4HepEm/G4HepEmData/include/G4HepEmElectronData.hh:81-354
 *   G4HepEmSBTableData        G4HepEm/G4HepEmData/include/G4HepEmSBTableData.hh:8-41
 *   G4HepEmGammaData          G4HepEm/G4HepEmData/include/G4HepEmGammaData.hh:20-78
 *
 * The reference declares a few members `const int` with in-class
 * initialisers (grid sizes); here they are plain `int` members at the same
 * offsets and the table builder writes the reference's values into them.
 * The offsets are checked member by member against the reference headers
 * by the static_asserts of oracle/ref/ref_harness.cc.
 *
 * Units follow Geant4's internal system: MeV, mm, 1/mm, MeV/mm.
 * The header has no includes beyond <stdbool.h> in C so that it can be
 * wrapped in a C++ namespace for the layout check.
 */
#ifndef G4HEPEM_AMD_DATA_H
#define G4HEPEM_AMD_DATA_H

#ifndef __cplusplus
#include <stdbool.h>
#endif

struct G4HepEmRegionParmeters {
  double fFinalRange;
  double fDRoverRange;
  double fLinELossLimit;
  double fMSCRangeFactor;
  double fMSCSafetyFactor;
  bool   fIsMSCMinimalStepLimit;
  bool   fIsELossFluctuation;
  bool   fIsMultipleStepsInMSCTrans;
  bool   fIsApplyCuts;
};

struct G4HepEmParameters {
  double fElectronTrackingCut;
  double fGammaTrackingCut;
  double fMinLossTableEnergy;
  double fMaxLossTableEnergy;
  int    fNumLossTableBins;
  double fElectronBremModelLim;
  bool   fIsMSCPositronCor;
  bool   fIsMSCDisplacement;
  int    fNumRegions;
  struct G4HepEmRegionParmeters* fParametersPerRegion;  /* [fNumRegions] */
};

struct G4HepEmMCCData {
  double fSecElProdCutE;
  double fSecPosProdCutE;
  double fSecGamProdCutE;
  double fLogSecGamCutE;
  int    fHepEmMatIndex;
  int    fG4MatCutIndex;
  int    fG4RegionIndex;
};

struct G4HepEmMatCutData {
  int    fNumG4MatCuts;
  int    fNumMatCutData;
  int*   fG4MCIndexToHepEmMCIndex;             /* [fNumG4MatCuts]  */
  struct G4HepEmMCCData* fMatCutData;          /* [fNumMatCutData] */
};

struct G4HepEmMatData {
  int     fG4MatIndex;
  int     fNumOfElement;
  int*    fElementVect;                        /* [fNumOfElement] Z values */
  double* fNumOfAtomsPerVolumeVect;            /* [fNumOfElement] 1/mm^3   */
  double  fDensity;
  double  fDensityCorFactor;
  double  fElectronDensity;
  double  fRadiationLength;
  double  fMeanExEnergy;
  int     fNumOfSandiaIntervals;
  double* fSandiaEnergies;                     /* [fNumOfSandiaIntervals]     */
  double* fSandiaCoefficients;                 /* [4 x fNumOfSandiaIntervals] */
  double  fZeff;
  double  fZeff23;
  double  fZeffSqrt;
  double  fUMSCPar;
  double  fUMSCStepMinPars[2];
  double  fUMSCTailCoeff[4];
  double  fUMSCThetaCoeff[2];
};

struct G4HepEmMaterialData {
  int    fNumG4Material;
  int    fNumMaterialData;
  int*   fG4MatIndexToHepEmMatIndex;           /* [fNumG4Material]   */
  struct G4HepEmMatData* fMaterialData;        /* [fNumMaterialData] */
};

struct G4HepEmElemData {
  double  fZet;
  double  fZet13;
  double  fZet23;
  double  fCoulomb;
  double  fLogZ;
  double  fZFactor1;
  double  fDeltaMaxLow;
  double  fDeltaMaxHigh;
  double  fILVarS1;
  double  fILVarS1Cond;
  int     fNumOfSandiaIntervals;
  double* fSandiaEnergies;                     /* [fNumOfSandiaIntervals]     */
  double* fSandiaCoefficients;                 /* [4 x fNumOfSandiaIntervals] */
  double  fKShellBindingEnergy;
};

struct G4HepEmElementData {
  int    fMaxZet;
  struct G4HepEmElemData* fElementData;        /* [fMaxZet+1], indexed by Z */
};

struct G4HepEmElectronData {
  int     fNumMatCuts;
  int     fNumMaterials;
  /* energy loss: per couple 5N values [R,R''] x N, [dEdx,dEdx''] x N, [InvR''] x N */
  int     fELossEnergyGridSize;
  double  fELossLogMinEkin;
  double  fELossEILDelta;
  double* fELossEnergyGrid;                    /* [N] */
  double* fELossData;                          /* [5 N fNumMatCuts] */
  /* restricted macroscopic cross sections: per couple ioni block then brem
     block, each [M, E(max), max, lnE0, 1/dlnE, (E, S, S'') x M]            */
  int     fResMacXSecNumData;
  int*    fResMacXSecStartIndexPerMatCut;      /* [fNumMatCuts] */
  double* fResMacXSecData;
  /* e-/e+ nuclear: 128 point grid per material, [S, S''] x 128 */
  int     fENucEnergyGridSize;
  double  fENucLogMinEkin;
  double  fENucEILDelta;
  double* fENucEnergyGrid;
  double* fENucMacXsecData;
  /* first transport cross section per material on the loss grid [S,S''] x N */
  double* fTr1MacXSecData;
  /* element selectors: [K, Q, lnE0, 1/dlnE, (E, P_0..P_{Q-2}) x K] */
  int     fElemSelectorIoniNumData;
  int*    fElemSelectorIoniStartIndexPerMatCut;
  double* fElemSelectorIoniData;
  int     fElemSelectorBremSBNumData;
  int*    fElemSelectorBremSBStartIndexPerMatCut;
  double* fElemSelectorBremSBData;
  int     fElemSelectorBremRBNumData;
  int*    fElemSelectorBremRBStartIndexPerMatCut;
  double* fElemSelectorBremRBData;
};

struct G4HepEmSBTableData {
  int     fMaxZet;                             /* 99 */
  int     fNumElEnergy;                        /* 65 */
  int     fNumKappa;                           /* 54 */
  double  fLogMinElEnergy;
  double  fILDeltaElEnergy;
  double  fElEnergyVect[65];
  double  fLElEnergyVect[65];
  double  fKappaVect[54];
  double  fLKappaVect[54];
  int     fNumHepEmMatCuts;
  int     fNumElemsInMatCuts;
  int*    fGammaCutIndxStartIndexPerMC;        /* [fNumHepEmMatCuts]   */
  int*    fGammaCutIndices;                    /* [fNumElemsInMatCuts] */
  int     fNumSBTableData;
  int     fSBTablesStartPerZ[121];
  double* fSBTableData;                        /* [fNumSBTableData] */
};

struct G4HepEmGammaData {
  int     fNumMaterials;
  int     fEGridSize0;                         /* 32  */
  int     fEGridSize1;                         /* 32  */
  int     fEGridSize2;                         /* 256 */
  int     fDataPerMat;
  int     fNumData0;
  int     fNumData1;
  double  fEMin0;
  double  fEMax0;
  double  fLogEMin0;
  double  fEILDelta0;
  double  fEMax1;
  double  fLogEMin1;
  double  fEILDelta1;
  double  fEMax2;
  double  fLogEMin2;
  double  fEILDelta2;
  double* fMacXsecData;                        /* [fNumMaterials x fDataPerMat] */
  int     fElemSelectorConvEgridSize;
  int     fElemSelectorConvNumData;
  double  fElemSelectorConvLogMinEkin;
  double  fElemSelectorConvEILDelta;
  int*    fElemSelectorConvStartIndexPerMat;   /* [fNumMaterials] */
  double* fElemSelectorConvEgrid;              /* [fElemSelectorConvEgridSize] */
  double* fElemSelectorConvData;               /* [fElemSelectorConvNumData]   */
};

struct G4HepEmData {
  struct G4HepEmMatCutData*   fTheMatCutData;
  struct G4HepEmMaterialData* fTheMaterialData;
  struct G4HepEmElementData*  fTheElementData;
  struct G4HepEmElectronData* fTheElectronData;
  struct G4HepEmElectronData* fThePositronData;
  struct G4HepEmSBTableData*  fTheSBTableData;
  struct G4HepEmGammaData*    fTheGammaData;
};

#endif /* G4HEPEM_AMD_DATA_H */
