// ref_harness.cc -- TEST INFRASTRUCTURE ONLY.
//
// Drives the reference's own G4HepEmRun code (compiled from the reference
// tree by oracle/Makefile into oracle/_ref/libhepem_ref.so) on this project's
// tables, to write golden vectors that pin the C restatement (hepem_oracle.c).
//
// The reference leaves G4HepEmRandomEngine::flat/flatArray to the consumer
// (G4HepEm/G4HepEmRun/include/G4HepEmRandomEngine.hh:22-47,
//  G4HepEm/G4HepEmRun/CMakeLists.txt:72-75); they are defined here on the
// same per-track Philox4x32-10 stream the oracle and the GPU use.
//
// The project's table structs are reinterpreted as the reference's: the
// static_asserts below check every member offset.

#include "G4HepEmData.hh"
#include "G4HepEmParameters.hh"
#include "G4HepEmMatCutData.hh"
#include "G4HepEmMaterialData.hh"
#include "G4HepEmElementData.hh"
#include "G4HepEmElectronData.hh"
#include "G4HepEmGammaData.hh"
#include "G4HepEmSBTableData.hh"

#include "G4HepEmElectronManager.hh"
#include "G4HepEmGammaManager.hh"
#include "G4HepEmElectronInteractionBrem.hh"
#include "G4HepEmGammaInteractionConversion.hh"
#include "G4HepEmElectronTrack.hh"
#include "G4HepEmGammaTrack.hh"
#include "G4HepEmTrack.hh"
#include "G4HepEmTLData.hh"
#include "G4HepEmRandomEngine.hh"

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace mine {
#include "g4hepem_amd.h"
}

// ---- layout check ----------------------------------------------------------------
#define SAME(T, f) static_assert(offsetof(mine::T, f) == offsetof(::T, f), #T "::" #f)
#define SIZE(T) static_assert(sizeof(mine::T) == sizeof(::T), "sizeof " #T)
SIZE(G4HepEmRegionParmeters); SAME(G4HepEmRegionParmeters, fFinalRange); SAME(G4HepEmRegionParmeters, fDRoverRange);
SAME(G4HepEmRegionParmeters, fLinELossLimit); SAME(G4HepEmRegionParmeters, fMSCRangeFactor);
SAME(G4HepEmRegionParmeters, fMSCSafetyFactor); SAME(G4HepEmRegionParmeters, fIsMSCMinimalStepLimit);
SAME(G4HepEmRegionParmeters, fIsELossFluctuation); SAME(G4HepEmRegionParmeters, fIsMultipleStepsInMSCTrans);
SAME(G4HepEmRegionParmeters, fIsApplyCuts);
SIZE(G4HepEmParameters); SAME(G4HepEmParameters, fElectronTrackingCut); SAME(G4HepEmParameters, fGammaTrackingCut);
SAME(G4HepEmParameters, fMinLossTableEnergy); SAME(G4HepEmParameters, fMaxLossTableEnergy);
SAME(G4HepEmParameters, fNumLossTableBins); SAME(G4HepEmParameters, fElectronBremModelLim);
SAME(G4HepEmParameters, fIsMSCPositronCor); SAME(G4HepEmParameters, fIsMSCDisplacement);
SAME(G4HepEmParameters, fNumRegions); SAME(G4HepEmParameters, fParametersPerRegion);
SIZE(G4HepEmMCCData); SAME(G4HepEmMCCData, fSecElProdCutE); SAME(G4HepEmMCCData, fSecPosProdCutE);
SAME(G4HepEmMCCData, fSecGamProdCutE); SAME(G4HepEmMCCData, fLogSecGamCutE); SAME(G4HepEmMCCData, fHepEmMatIndex);
SAME(G4HepEmMCCData, fG4MatCutIndex); SAME(G4HepEmMCCData, fG4RegionIndex);
SIZE(G4HepEmMatCutData); SAME(G4HepEmMatCutData, fNumG4MatCuts); SAME(G4HepEmMatCutData, fNumMatCutData);
SAME(G4HepEmMatCutData, fG4MCIndexToHepEmMCIndex); SAME(G4HepEmMatCutData, fMatCutData);
SIZE(G4HepEmMatData); SAME(G4HepEmMatData, fG4MatIndex); SAME(G4HepEmMatData, fNumOfElement); SAME(G4HepEmMatData, fElementVect);
SAME(G4HepEmMatData, fNumOfAtomsPerVolumeVect); SAME(G4HepEmMatData, fDensity); SAME(G4HepEmMatData, fDensityCorFactor);
SAME(G4HepEmMatData, fElectronDensity); SAME(G4HepEmMatData, fRadiationLength); SAME(G4HepEmMatData, fMeanExEnergy);
SAME(G4HepEmMatData, fNumOfSandiaIntervals); SAME(G4HepEmMatData, fSandiaEnergies); SAME(G4HepEmMatData, fSandiaCoefficients);
SAME(G4HepEmMatData, fZeff); SAME(G4HepEmMatData, fZeff23); SAME(G4HepEmMatData, fZeffSqrt); SAME(G4HepEmMatData, fUMSCPar);
SAME(G4HepEmMatData, fUMSCStepMinPars); SAME(G4HepEmMatData, fUMSCTailCoeff); SAME(G4HepEmMatData, fUMSCThetaCoeff);
SIZE(G4HepEmMaterialData); SAME(G4HepEmMaterialData, fNumG4Material); SAME(G4HepEmMaterialData, fNumMaterialData);
SAME(G4HepEmMaterialData, fG4MatIndexToHepEmMatIndex); SAME(G4HepEmMaterialData, fMaterialData);
SIZE(G4HepEmElemData); SAME(G4HepEmElemData, fZet); SAME(G4HepEmElemData, fZet13); SAME(G4HepEmElemData, fZet23);
SAME(G4HepEmElemData, fCoulomb); SAME(G4HepEmElemData, fLogZ); SAME(G4HepEmElemData, fZFactor1);
SAME(G4HepEmElemData, fDeltaMaxLow); SAME(G4HepEmElemData, fDeltaMaxHigh); SAME(G4HepEmElemData, fILVarS1);
SAME(G4HepEmElemData, fILVarS1Cond); SAME(G4HepEmElemData, fNumOfSandiaIntervals); SAME(G4HepEmElemData, fSandiaEnergies);
SAME(G4HepEmElemData, fSandiaCoefficients); SAME(G4HepEmElemData, fKShellBindingEnergy);
SIZE(G4HepEmElementData); SAME(G4HepEmElementData, fMaxZet); SAME(G4HepEmElementData, fElementData);
SIZE(G4HepEmElectronData); SAME(G4HepEmElectronData, fNumMatCuts); SAME(G4HepEmElectronData, fNumMaterials);
SAME(G4HepEmElectronData, fELossEnergyGridSize); SAME(G4HepEmElectronData, fELossLogMinEkin); SAME(G4HepEmElectronData, fELossEILDelta);
SAME(G4HepEmElectronData, fELossEnergyGrid); SAME(G4HepEmElectronData, fELossData); SAME(G4HepEmElectronData, fResMacXSecNumData);
SAME(G4HepEmElectronData, fResMacXSecStartIndexPerMatCut); SAME(G4HepEmElectronData, fResMacXSecData);
SAME(G4HepEmElectronData, fENucEnergyGridSize); SAME(G4HepEmElectronData, fENucLogMinEkin); SAME(G4HepEmElectronData, fENucEILDelta);
SAME(G4HepEmElectronData, fENucEnergyGrid); SAME(G4HepEmElectronData, fENucMacXsecData); SAME(G4HepEmElectronData, fTr1MacXSecData);
SAME(G4HepEmElectronData, fElemSelectorIoniNumData); SAME(G4HepEmElectronData, fElemSelectorIoniStartIndexPerMatCut);
SAME(G4HepEmElectronData, fElemSelectorIoniData); SAME(G4HepEmElectronData, fElemSelectorBremSBNumData);
SAME(G4HepEmElectronData, fElemSelectorBremSBStartIndexPerMatCut); SAME(G4HepEmElectronData, fElemSelectorBremSBData);
SAME(G4HepEmElectronData, fElemSelectorBremRBNumData); SAME(G4HepEmElectronData, fElemSelectorBremRBStartIndexPerMatCut);
SAME(G4HepEmElectronData, fElemSelectorBremRBData);
SIZE(G4HepEmSBTableData); SAME(G4HepEmSBTableData, fMaxZet); SAME(G4HepEmSBTableData, fNumElEnergy); SAME(G4HepEmSBTableData, fNumKappa);
SAME(G4HepEmSBTableData, fLogMinElEnergy); SAME(G4HepEmSBTableData, fILDeltaElEnergy); SAME(G4HepEmSBTableData, fElEnergyVect);
SAME(G4HepEmSBTableData, fLElEnergyVect); SAME(G4HepEmSBTableData, fKappaVect); SAME(G4HepEmSBTableData, fLKappaVect);
SAME(G4HepEmSBTableData, fNumHepEmMatCuts); SAME(G4HepEmSBTableData, fNumElemsInMatCuts);
SAME(G4HepEmSBTableData, fGammaCutIndxStartIndexPerMC); SAME(G4HepEmSBTableData, fGammaCutIndices);
SAME(G4HepEmSBTableData, fNumSBTableData); SAME(G4HepEmSBTableData, fSBTablesStartPerZ); SAME(G4HepEmSBTableData, fSBTableData);
SIZE(G4HepEmGammaData); SAME(G4HepEmGammaData, fNumMaterials); SAME(G4HepEmGammaData, fEGridSize0); SAME(G4HepEmGammaData, fEGridSize1);
SAME(G4HepEmGammaData, fEGridSize2); SAME(G4HepEmGammaData, fDataPerMat); SAME(G4HepEmGammaData, fNumData0);
SAME(G4HepEmGammaData, fNumData1); SAME(G4HepEmGammaData, fEMin0); SAME(G4HepEmGammaData, fEMax0); SAME(G4HepEmGammaData, fLogEMin0);
SAME(G4HepEmGammaData, fEILDelta0); SAME(G4HepEmGammaData, fEMax1); SAME(G4HepEmGammaData, fLogEMin1); SAME(G4HepEmGammaData, fEILDelta1);
SAME(G4HepEmGammaData, fEMax2); SAME(G4HepEmGammaData, fLogEMin2); SAME(G4HepEmGammaData, fEILDelta2);
SAME(G4HepEmGammaData, fMacXsecData); SAME(G4HepEmGammaData, fElemSelectorConvEgridSize); SAME(G4HepEmGammaData, fElemSelectorConvNumData);
SAME(G4HepEmGammaData, fElemSelectorConvLogMinEkin); SAME(G4HepEmGammaData, fElemSelectorConvEILDelta);
SAME(G4HepEmGammaData, fElemSelectorConvStartIndexPerMat); SAME(G4HepEmGammaData, fElemSelectorConvEgrid);
SAME(G4HepEmGammaData, fElemSelectorConvData);
SIZE(G4HepEmData); SAME(G4HepEmData, fTheMatCutData); SAME(G4HepEmData, fTheMaterialData); SAME(G4HepEmData, fTheElementData);
SAME(G4HepEmData, fTheElectronData); SAME(G4HepEmData, fThePositronData); SAME(G4HepEmData, fTheSBTableData);
SAME(G4HepEmData, fTheGammaData);
#undef SAME
#undef SIZE

// ---- random engine binding ------------------------------------------------------------
namespace {
struct Stream { uint32_t key[2]; uint32_t id; uint32_t hi; uint32_t ctr; };

inline void Philox(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3], k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = uint64_t(0xD2511F53u) * c0, p1 = uint64_t(0xCD9E8D57u) * c2;
    const uint32_t nc0 = uint32_t(p1 >> 32) ^ c1 ^ k0, nc2 = uint32_t(p0 >> 32) ^ c3 ^ k1;
    c1 = uint32_t(p1); c3 = uint32_t(p0); c0 = nc0; c2 = nc2;
    k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}
inline double Draw(Stream* s) {
  const uint32_t d = s->ctr++;
  const uint32_t c[4] = {d >> 1, s->id, s->hi, 0u};
  uint32_t o[4];
  Philox(c, s->key, o);
  const int j = int(d & 1u);
  const uint64_t w = (uint64_t(o[2 * j + 1]) << 32) | uint64_t(o[2 * j]);
  return (double(w >> 12) + 0.5) * 2.220446049250313080847e-16;
}
// G4HepEmRandomEngine's private state {fObject, fIsGauss, fGauss} (G4HepEmRandomEngine.hh:100-104)
struct EngineMirror { void* obj; bool isGauss; double gauss; };
static_assert(sizeof(EngineMirror) == sizeof(G4HepEmRandomEngine), "engine layout");
}  // namespace

double G4HepEmRandomEngine::flat() { return Draw(static_cast<Stream*>(fObject)); }
void G4HepEmRandomEngine::flatArray(const int size, double* vect) {
  for (int i = 0; i < size; ++i) vect[i] = Draw(static_cast<Stream*>(fObject));
}

// ---- drivers ----------------------------------------------------------------------------
namespace {
inline G4HepEmData* R(const mine::G4HepEmData* d) { return reinterpret_cast<G4HepEmData*>(const_cast<mine::G4HepEmData*>(d)); }
inline G4HepEmParameters* R(const mine::G4HepEmParameters* p) { return reinterpret_cast<G4HepEmParameters*>(const_cast<mine::G4HepEmParameters*>(p)); }

// the caller's geometry between HowFar and Perform: a step longer than
// geom_limit[i] is cut there and ends on a boundary, any other does not
void ApplyGeometry(const double* geom_limit, int i, G4HepEmTrack* t) {
  if (!geom_limit) return;
  if (geom_limit[i] < t->GetGStepLength()) {
    t->SetGStepLength(geom_limit[i]);
    t->SetOnBoundary(true);
  } else {
    t->SetOnBoundary(false);
  }
}

void StepRange(G4HepEmData* d, G4HepEmParameters* p, uint64_t seed, const mine::hepem_tracks* in, mine::hepem_tracks* out,
               mine::hepem_step_out* res, const double* geom_limit, int begin, int end) {
  const int n = in->n;
  G4HepEmTLData tl;
  for (int i = begin; i < end; ++i) {
    Stream st{{uint32_t(seed), uint32_t(seed >> 32)}, uint32_t(in->id[i]), in->rng_stream ? in->rng_stream[i] : 0u,
              in->rng_ctr[i]};
    G4HepEmRandomEngine eng(&st);
    EngineMirror* em = reinterpret_cast<EngineMirror*>(&eng);
    em->isGauss = in->rng_has_gauss[i] != 0;
    em->gauss = in->rng_gauss[i];
    tl.SetRandomEngine(&eng);
    tl.ResetNumSecondaryElectronTrack();
    tl.ResetNumSecondaryGammaTrack();
    tl.GetSecondaryElectronTrack(0)->ReSet();
    tl.GetSecondaryElectronTrack(1)->ReSet();
    const int ptype = in->ptype[i];
    G4HepEmTrack* t;
    double displ[3] = {0., 0., 0.}, trueStep, geomStep;
    if (ptype == mine::HEPEM_GAMMA) {
      G4HepEmGammaTrack* gt = tl.GetPrimaryGammaTrack();
      gt->ReSet();
      t = gt->GetTrack();
    } else {
      G4HepEmElectronTrack* et = tl.GetPrimaryElectronTrack();
      et->ReSet();
      t = et->GetTrack();
      t->SetCharge(ptype == mine::HEPEM_ELECTRON ? -1.0 : 1.0);
      G4HepEmMSCTrackData* m = et->GetMSCTrackData();
      m->fInitialRange = in->msc_init_range[i];
      m->fDynamicRangeFactor = in->msc_dyn_range_factor[i];
      m->fTlimitMin = in->msc_tlimit_min[i];
      m->fIsFirstStep = in->msc_first_step[i] != 0;
    }
    t->SetEKin(in->ekin[i]);
    t->SetDirection(in->dir[i], in->dir[n + i], in->dir[2 * n + i]);
    t->SetMCIndex(in->imc[i]);
    t->SetSafety(in->safety[i]);
    t->SetOnBoundary(in->on_boundary[i] != 0);
    for (int k = 0; k < 4; ++k) t->SetNumIALeft(in->num_ia_left[k * n + i], k);
    if (ptype == mine::HEPEM_GAMMA) {
      G4HepEmGammaManager::HowFar(d, p, &tl);
      ApplyGeometry(geom_limit, i, t);
      if (t->GetOnBoundary()) {
        G4HepEmGammaManager::UpdateNumIALeft(t);
        t->SetEnergyDeposit(0.0);
      } else {
        G4HepEmGammaManager::SelectInteraction(d, &tl);
        if (t->GetWinnerProcessIndex() != 3) G4HepEmGammaManager::Perform(d, p, &tl);
        else t->SetEnergyDeposit(0.0);
      }
      trueStep = geomStep = t->GetGStepLength();
    } else {
      G4HepEmElectronManager::HowFar(d, p, &tl);
      ApplyGeometry(geom_limit, i, t);
      G4HepEmElectronManager::Perform(d, p, &tl);
      G4HepEmElectronTrack* et = tl.GetPrimaryElectronTrack();
      G4HepEmMSCTrackData* m = et->GetMSCTrackData();
      trueStep = et->GetPStepLength();
      geomStep = t->GetGStepLength();
      for (int c = 0; c < 3; ++c) displ[c] = m->GetDisplacement()[c];
      out->msc_init_range[i] = m->fInitialRange;
      out->msc_dyn_range_factor[i] = m->fDynamicRangeFactor;
      out->msc_tlimit_min[i] = m->fTlimitMin;
      out->msc_first_step[i] = m->fIsFirstStep ? 1 : 0;
    }
    out->ekin[i] = t->GetEKin();
    for (int c = 0; c < 3; ++c) out->dir[c * n + i] = t->GetDirection()[c];
    for (int k = 0; k < 4; ++k) out->num_ia_left[k * n + i] = t->GetNumIALeft(k);
    out->rng_ctr[i] = st.ctr;
    out->rng_gauss[i] = em->gauss;
    out->rng_has_gauss[i] = em->isGauss ? 1 : 0;
    if (out != in) {
      out->ptype[i] = in->ptype[i];
      out->id[i] = in->id[i];
      if (out->rng_stream) out->rng_stream[i] = in->rng_stream ? in->rng_stream[i] : 0u;
      out->imc[i] = in->imc[i];
      out->safety[i] = in->safety[i];
    }
    out->on_boundary[i] = t->GetOnBoundary() ? 1 : 0;
    res->true_step[i] = trueStep;
    res->geom_step[i] = geomStep;
    res->edep[i] = t->GetEnergyDeposit();
    for (int c = 0; c < 3; ++c) res->displacement[c * n + i] = displ[c];
    res->winner[i] = t->GetWinnerProcessIndex();
    res->stopped[i] = t->GetEKin() <= 0.0 ? 1 : 0;
    int ns = 0;
    const int nel = int(tl.GetNumSecondaryElectronTrack()), ngm = int(tl.GetNumSecondaryGammaTrack());
    for (int k = 0; k < nel + ngm && ns < 2; ++k, ++ns) {
      G4HepEmTrack* s = k < nel ? tl.GetSecondaryElectronTrack(k)->GetTrack() : tl.GetSecondaryGammaTrack(k - nel)->GetTrack();
      res->sec_ptype[ns * n + i] = k < nel ? (s->GetCharge() < 0.0 ? mine::HEPEM_ELECTRON : mine::HEPEM_POSITRON) : mine::HEPEM_GAMMA;
      res->sec_ekin[ns * n + i] = s->GetEKin();
      for (int c = 0; c < 3; ++c) res->sec_dir[(c * 2 + ns) * n + i] = s->GetDirection()[c];
    }
    res->nsec[i] = nel + ngm;
    for (int s = ns; s < 2; ++s) {
      res->sec_ptype[s * n + i] = -1;
      res->sec_ekin[s * n + i] = 0.0;
      for (int c = 0; c < 3; ++c) res->sec_dir[(c * 2 + s) * n + i] = 0.0;
    }
  }
}
}  // namespace

extern "C" {

int ref_layout_ok() { return 1; }

int ref_lookup_eloss_xsec(const mine::G4HepEmData* dd, int ptype, int n, const int32_t* imc, const double* ekin,
                          const double* lekin, double* out) {
  G4HepEmData* d = R(dd);
  const G4HepEmElectronData* e = ptype == 0 ? d->fTheElectronData : d->fThePositronData;
  for (int i = 0; i < n; ++i) {
    const double r = G4HepEmElectronManager::GetRestRange(e, imc[i], ekin[i], lekin[i]);
    out[i] = G4HepEmElectronManager::GetRestDEDX(e, imc[i], ekin[i], lekin[i]);
    out[n + i] = r;
    out[2 * n + i] = G4HepEmElectronManager::GetInvRange(e, imc[i], r);
    out[3 * n + i] = G4HepEmElectronManager::GetRestMacXSec(e, imc[i], ekin[i], lekin[i], true);
    out[4 * n + i] = G4HepEmElectronManager::GetRestMacXSec(e, imc[i], ekin[i], lekin[i], false);
  }
  return 0;
}

int ref_probe_electron(const mine::G4HepEmData* dd, int ptype, int which, int n, const int32_t* imc, const double* ekin,
                       const double* lekin, double* out) {
  G4HepEmData* d = R(dd);
  const G4HepEmElectronData* e = ptype == 0 ? d->fTheElectronData : d->fThePositronData;
  for (int i = 0; i < n; ++i) {
    const int imat = d->fTheMatCutData->fMatCutData[imc[i]].fHepEmMatIndex;
    const double ne = d->fTheMaterialData->fMaterialData[imat].fElectronDensity;
    switch (which) {
      case 0: out[i] = G4HepEmElectronManager::GetRestMacXSecForStepping(e, imc[i], ekin[i], lekin[i], true); break;
      case 1: out[i] = G4HepEmElectronManager::GetRestMacXSecForStepping(e, imc[i], ekin[i], lekin[i], false); break;
      case 2: out[i] = G4HepEmElectronManager::GetTransportMFP(e, imat, ekin[i], lekin[i]); break;
      case 3: out[i] = G4HepEmElectronManager::GetMacXSecNuclear(e, imat, ekin[i], lekin[i]); break;
      case 4: out[i] = G4HepEmElectronManager::ComputeMacXsecAnnihilation(ekin[i], ne); break;
      case 5: out[i] = G4HepEmElectronManager::ComputeMacXsecAnnihilationForStepping(ekin[i], ne); break;
      default: return -1;
    }
  }
  return 0;
}

int ref_probe_selector(const mine::G4HepEmData* dd, int ptype, int which, int n, const int32_t* imc, const double* ekin,
                       const double* lekin, const double* urnd, int32_t* out) {
  G4HepEmData* d = R(dd);
  const G4HepEmElectronData* e = ptype == 0 ? d->fTheElectronData : d->fThePositronData;
  for (int i = 0; i < n; ++i) {
    const int imat = d->fTheMatCutData->fMatCutData[imc[i]].fHepEmMatIndex;
    if (d->fTheMaterialData->fMaterialData[imat].fNumOfElement < 2) { out[i] = 0; continue; }
    switch (which) {
      case 0: out[i] = G4HepEmElectronInteractionBrem::SelectTargetAtom(e, imc[i], ekin[i], lekin[i], urnd[i], true); break;
      case 1: out[i] = G4HepEmElectronInteractionBrem::SelectTargetAtom(e, imc[i], ekin[i], lekin[i], urnd[i], false); break;
      case 2: out[i] = G4HepEmGammaInteractionConversion::SelectTargetAtom(d->fTheGammaData, imat, ekin[i], lekin[i], urnd[i]); break;
      default: return -1;
    }
  }
  return 0;
}

int ref_probe_gamma(const mine::G4HepEmData* dd, int n, const int32_t* imc, const double* ekin, const double* lekin,
                    const double* urnd, double* out_xs, double* out_pe, int32_t* out_pid) {
  G4HepEmData* d = R(dd);
  for (int i = 0; i < n; ++i) {
    G4HepEmGammaTrack g;
    g.ReSet();
    G4HepEmTrack* t = g.GetTrack();
    t->SetEKin(ekin[i], lekin[i]);
    t->SetMCIndex(imc[i]);
    const double tot = G4HepEmGammaManager::GetTotalMacXSec(d, &g);
    t->SetMFP(tot > 0. ? 1. / tot : 1.0E+20, 0);
    out_xs[i] = tot;
    G4HepEmGammaManager::SampleInteraction(d, &g, urnd[i]);
    out_pe[i] = g.GetPEmxSec();
    out_pid[i] = t->GetWinnerProcessIndex();
  }
  return 0;
}

int ref_step_geom(const mine::G4HepEmData* dd, const mine::G4HepEmParameters* pp, uint64_t seed, const mine::hepem_tracks* in,
                  mine::hepem_tracks* out, mine::hepem_step_out* res, const double* geom_limit, int nthreads);

int ref_step(const mine::G4HepEmData* dd, const mine::G4HepEmParameters* pp, uint64_t seed, const mine::hepem_tracks* in,
             mine::hepem_tracks* out, mine::hepem_step_out* res, int nthreads) {
  return ref_step_geom(dd, pp, seed, in, out, res, nullptr, nthreads);
}

int ref_step_geom(const mine::G4HepEmData* dd, const mine::G4HepEmParameters* pp, uint64_t seed, const mine::hepem_tracks* in,
                  mine::hepem_tracks* out, mine::hepem_step_out* res, const double* geom_limit, int nthreads) {
  G4HepEmData* d = R(dd);
  G4HepEmParameters* p = R(pp);
  const int n = in->n;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > n) nthreads = n > 0 ? n : 1;
  std::vector<std::thread> th;
  const int chunk = (n + nthreads - 1) / nthreads;
  for (int k = 0; k < nthreads; ++k) {
    const int b = std::min(n, k * chunk), e = std::min(n, (k + 1) * chunk);
    th.emplace_back(StepRange, d, p, seed, in, out, res, geom_limit, b, e);
  }
  for (auto& t : th) t.join();
  return 0;
}

}  // extern "C"
