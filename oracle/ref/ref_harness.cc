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
#include "G4HepEmPositronInteractionAnnihilation.hh"
#include "G4HepEmMath.hh"

#include <cstddef>
#include <cstdint>
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

namespace mine {
#include "g4hepem_amd.h"
#include "g4hepem_amd_shower.h"
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
// One e-/e+ step as G4HepEmTrackingManager::TrackElectron takes it
// (G4HepEm/G4HepEm/src/G4HepEmTrackingManager.cc:408-635), driving the
// reference's own G4HepEmElectronManager functions in that order, with the
// caller's geometry a slab lo <= x <= hi across the x axis (the navigator's
// MakeStep: the straight move to the next plane at most; ComputeSafety: the
// distance to the nearer plane). Crossing lo (exit_lo) or hi (exit_hi) leaves
// the world. The reference's sub-step loop is unbounded; it is cut at
// kMaxSubsteps here as in the library.
constexpr int kMaxSubsteps = 65536;
struct SlabGeo {
  double x, lo, hi;
  bool exit_lo, exit_hi;
  int side = 0;
  bool left = false;
  double geom_path = 0.0, true_path = 0.0;
  int nsub = 0;
};

bool ElectronSlabStep(G4HepEmData* d, G4HepEmParameters* p, G4HepEmTLData* tl, SlabGeo& g) {
  G4HepEmElectronTrack* theElTrack = tl->GetPrimaryElectronTrack();
  G4HepEmTrack* thePrimaryTrack = theElTrack->GetTrack();
  G4HepEmRandomEngine* rnge = tl->GetRNGEngine();
  const bool isElectron = thePrimaryTrack->GetCharge() < 0.0;
  double x = g.x;
  const double preStepEkin = thePrimaryTrack->GetEKin();
  const double preStepLogEkin = thePrimaryTrack->GetLogEKin();
  const bool preStepOnBoundary = thePrimaryTrack->GetOnBoundary();
  const double preSafety = preStepOnBoundary ? 0. : std::max(0.0, std::min(x - g.lo, g.hi - x));
  thePrimaryTrack->SetSafety(preSafety);
  const int indxRegion = d->fTheMatCutData->fMatCutData[thePrimaryTrack->GetMCIndex()].fG4RegionIndex;
  bool continueStepping = p->fParametersPerRegion[indxRegion].fIsMultipleStepsInMSCTrans;
  for (int ip = 0; ip < 4; ++ip) {
    if (thePrimaryTrack->GetNumIALeft(ip) <= 0.) {
      thePrimaryTrack->SetNumIALeft(-G4HepEmLog(rnge->flat()), ip);
    }
  }
  G4HepEmElectronManager::HowFarToDiscreteInteraction(d, p, theElTrack);
  const int iDProc = thePrimaryTrack->GetWinnerProcessIndex();
  double stepLimitLeft = theElTrack->GetPStepLength();
  double totalTruePathLength = 0, totalEloss = 0;
  bool stopped = false;
  theElTrack->SavePreStepEKin();
  int nsub = 0;
  do {
    ++nsub;
    G4HepEmElectronManager::HowFarToMSC(d, p, theElTrack, rnge);
    if (thePrimaryTrack->GetWinnerProcessIndex() != -2) continueStepping = false;
    const double physicalStep = thePrimaryTrack->GetGStepLength();
    const double ux = thePrimaryTrack->GetDirection()[0];
    double geometryStep = std::numeric_limits<double>::max();
    int sd = 0;
    if (ux > 0.0) { geometryStep = (g.hi - x) / ux; sd = 1; }
    else if (ux < 0.0) { geometryStep = (g.lo - x) / ux; sd = -1; }
    const bool geometryLimitedStep = geometryStep < physicalStep;
    const double finalStep = geometryLimitedStep ? geometryStep : physicalStep;
    x = geometryLimitedStep ? (sd > 0 ? g.hi : g.lo) : std::min(std::max(x + finalStep * ux, g.lo), g.hi);
    g.geom_path += finalStep;
    if (geometryLimitedStep) {
      continueStepping = false;
      g.side = sd;
      if ((sd > 0 && g.exit_hi) || (sd < 0 && g.exit_lo)) {
        g.left = true;
        break;
      }
    }
    const bool postStepOnBoundary = geometryLimitedStep;
    thePrimaryTrack->SetGStepLength(finalStep);
    thePrimaryTrack->SetOnBoundary(postStepOnBoundary);
    if (finalStep > 0) {
      do {
        G4HepEmElectronManager::UpdatePStepLength(theElTrack);
        const double pStepLength = theElTrack->GetPStepLength();
        totalTruePathLength += pStepLength;
        if (pStepLength <= 0.0) break;
        G4HepEmElectronManager::UpdateNumIALeft(theElTrack);
        stopped = G4HepEmElectronManager::ApplyMeanEnergyLoss(d, p, theElTrack);
        totalEloss += thePrimaryTrack->GetEnergyDeposit();
        if (stopped) {
          continueStepping = false;
          break;
        }
        G4HepEmElectronManager::SampleMSC(d, p, theElTrack, rnge);
        if (!postStepOnBoundary) {
          const double* displacement = theElTrack->GetMSCTrackData()->GetDisplacement();
          const double dLength2 = displacement[0] * displacement[0] + displacement[1] * displacement[1] +
                                  displacement[2] * displacement[2];
          const double kGeomMinLength = 5.0e-8;
          if (dLength2 > kGeomMinLength * kGeomMinLength) {
            const double dispR = std::sqrt(dLength2);
            const double postSafety = 0.99 * std::max(0.0, std::min(x - g.lo, g.hi - x));
            if (postSafety > 0.0 && dispR <= postSafety) {
              x += displacement[0];
            } else if (dispR < postSafety) {
              x += displacement[0];
            } else if (postSafety > kGeomMinLength) {
              x += displacement[0] * (postSafety / dispR);
            }
          }
        }
      } while (0);
      if (continueStepping) {
        thePrimaryTrack->SetEnergyDeposit(0);
        thePrimaryTrack->SetWinnerProcessIndex(iDProc);
        theElTrack->SavePreStepEKin();
        const double pStepLength = theElTrack->GetPStepLength();
        stepLimitLeft -= pStepLength;
        theElTrack->SetPStepLength(stepLimitLeft);
        thePrimaryTrack->SetGStepLength(stepLimitLeft);
        theElTrack->SetRange(theElTrack->GetRange() - pStepLength);
      }
    }
  } while (continueStepping && nsub < kMaxSubsteps);
  g.x = x;
  g.nsub = nsub;
  g.true_path = totalTruePathLength;
  thePrimaryTrack->SetEnergyDeposit(totalEloss);
  if (!stopped) {
    theElTrack->SetPreStepEKin(preStepEkin, preStepLogEkin);
    stopped = G4HepEmElectronManager::SampleLossFluctuations(d, p, theElTrack, rnge);
  }
  if (stopped) {
    if (!isElectron) G4HepEmPositronInteractionAnnihilation::Perform(tl, true);
  } else if (!g.left) {
    // PerformDiscrete: no discrete interaction on a boundary; winner 3 resets
    // its interaction length and draws the delta test, as TrackElectron does
    // with a lepto-nuclear process attached (:651-664)
    G4HepEmElectronManager::PerformDiscrete(d, p, tl);
  }
  return stopped;
}

void ElectronSlabRange(G4HepEmData* d, G4HepEmParameters* p, uint64_t seed, mine::hepem_tracks* T,
                       mine::hepem_step_out* res, double* xs, const double* lo, const double* hi, const uint8_t* exitf,
                       int8_t* side, uint8_t* left, int32_t* nsub, int begin, int end) {
  const int n = T->n;
  G4HepEmTLData tl;
  for (int i = begin; i < end; ++i) {
    Stream st{{uint32_t(seed), uint32_t(seed >> 32)}, uint32_t(T->id[i]), T->rng_stream ? T->rng_stream[i] : 0u,
              T->rng_ctr[i]};
    G4HepEmRandomEngine eng(&st);
    EngineMirror* em = reinterpret_cast<EngineMirror*>(&eng);
    em->isGauss = T->rng_has_gauss[i] != 0;
    em->gauss = T->rng_gauss[i];
    tl.SetRandomEngine(&eng);
    tl.ResetNumSecondaryElectronTrack();
    tl.ResetNumSecondaryGammaTrack();
    G4HepEmElectronTrack* et = tl.GetPrimaryElectronTrack();
    et->ReSet();
    G4HepEmTrack* t = et->GetTrack();
    t->SetCharge(T->ptype[i] == mine::HEPEM_ELECTRON ? -1.0 : 1.0);
    G4HepEmMSCTrackData* m = et->GetMSCTrackData();
    m->fInitialRange = T->msc_init_range[i];
    m->fDynamicRangeFactor = T->msc_dyn_range_factor[i];
    m->fTlimitMin = T->msc_tlimit_min[i];
    m->fIsFirstStep = T->msc_first_step[i] != 0;
    t->SetEKin(T->ekin[i]);
    t->SetDirection(T->dir[i], T->dir[n + i], T->dir[2 * n + i]);
    t->SetMCIndex(T->imc[i]);
    t->SetOnBoundary(T->on_boundary[i] != 0);
    for (int k = 0; k < 4; ++k) t->SetNumIALeft(T->num_ia_left[k * n + i], k);
    SlabGeo g;
    g.x = xs[i];
    g.lo = lo[i];
    g.hi = hi[i];
    g.exit_lo = (exitf[i] & 1) != 0;
    g.exit_hi = (exitf[i] & 2) != 0;
    const bool stopped = ElectronSlabStep(d, p, &tl, g);
    T->ekin[i] = t->GetEKin();
    for (int c = 0; c < 3; ++c) T->dir[c * n + i] = t->GetDirection()[c];
    for (int k = 0; k < 4; ++k) T->num_ia_left[k * n + i] = t->GetNumIALeft(k);
    T->safety[i] = t->GetSafety();
    T->on_boundary[i] = t->GetOnBoundary() ? 1 : 0;
    T->msc_init_range[i] = m->fInitialRange;
    T->msc_dyn_range_factor[i] = m->fDynamicRangeFactor;
    T->msc_tlimit_min[i] = m->fTlimitMin;
    T->msc_first_step[i] = m->fIsFirstStep ? 1 : 0;
    T->rng_ctr[i] = st.ctr;
    T->rng_gauss[i] = em->gauss;
    T->rng_has_gauss[i] = em->isGauss ? 1 : 0;
    xs[i] = g.x;
    side[i] = int8_t(g.side);
    left[i] = g.left ? 1 : 0;
    nsub[i] = g.nsub;
    res->true_step[i] = g.true_path;
    res->geom_step[i] = g.geom_path;
    res->edep[i] = t->GetEnergyDeposit();
    for (int c = 0; c < 3; ++c) res->displacement[c * n + i] = 0.0;
    res->winner[i] = t->GetWinnerProcessIndex();
    res->stopped[i] = t->GetEKin() <= 0.0 ? 1 : 0;  // stopped or killed in the step
    (void)stopped;
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
// ---- showers in a stack of slabs, stepped by the reference's own code -------------------
// The event loop of G4HepEmTrackingManager (TrackElectron above; TrackGamma,
// G4HepEm/G4HepEm/src/G4HepEmTrackingManager.cc:893-1145, without Woodcock
// tracking; StackSecondaries, :1274-1313) over a 1D stack of slabs: each event
// starts with one primary at x = 0 moving along +x, tracks are followed
// depth-first from a stack, a track's stream selector is the splitmix64 mix of
// its parent's with (step number, secondary slot) and an event's is that of
// its index, deposits are summed per (event, slab) in 1e-9 MeV fixed point.
// Energy leaving the stack (a track's kinetic energy, an e+'s annihilation
// energy and what its last step created) is counted per event.
constexpr double kShFix = 1.0e9;
constexpr double kMe = 0.51099895000;

uint64_t ShMix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct ShTrack {
  int ptype = 0;
  uint64_t stream = 0;
  int imc = 0;
  double ekin = 0.0, dir[3] = {1.0, 0.0, 0.0};
  bool onBoundary = false;
  double nial[4] = {-1.0, -1.0, -1.0, -1.0};
  double mscInitRange = 1.0e+21, mscDynRangeFactor = 0.04, mscTlimitMin = 1.0E-7;
  bool mscFirstStep = true;
  uint32_t ctr = 0;
  bool hasGauss = false;
  double gauss = 0.0;
  double x = 0.0;
  int layer = 0, nstep = 0;
};

struct ShRun {
  G4HepEmData* d;
  G4HepEmParameters* p;
  uint64_t seed;
  const mine::hepem_slab_setup* su;
  std::vector<double> plane;
  int64_t* dep;
  int64_t* leak;
  std::atomic<int64_t> steps{0}, tracks{0};
};

void ShEvent(ShRun& run, int ev) {
  G4HepEmData* d = run.d;
  G4HepEmParameters* p = run.p;
  const int nabs = run.su->nabs;
  const int nbins = run.su->nlayers * nabs;
  const double* P = run.plane.data();
  int64_t* dep = run.dep + (size_t)ev * nbins;
  int64_t steps = 0, tracks = 0;
  std::vector<ShTrack> stack;
  {
    ShTrack t;
    t.ptype = run.su->primary;
    t.stream = ShMix64(0x5EED000000000000ull ^ (uint64_t)ev);
    t.imc = run.su->imc[0];
    t.ekin = run.su->ekin0;
    t.onBoundary = true;
    stack.push_back(t);
    ++tracks;
  }
  G4HepEmTLData tl;
  while (!stack.empty()) {
    ShTrack t = stack.back();
    stack.pop_back();
    while (true) {
      int l = t.layer;
      const int l0 = l;
      const double lo = P[l], hi = P[l + 1];
      Stream st{{uint32_t(run.seed), uint32_t(run.seed >> 32)}, uint32_t(t.stream), uint32_t(t.stream >> 32), t.ctr};
      G4HepEmRandomEngine eng(&st);
      EngineMirror* em = reinterpret_cast<EngineMirror*>(&eng);
      em->isGauss = t.hasGauss;
      em->gauss = t.gauss;
      tl.SetRandomEngine(&eng);
      tl.ResetNumSecondaryElectronTrack();
      tl.ResetNumSecondaryGammaTrack();
      double x, edep;
      int sd = 0;
      bool onb, left = false;
      G4HepEmTrack* pt;
      if (t.ptype != mine::HEPEM_GAMMA) {
        G4HepEmElectronTrack* et = tl.GetPrimaryElectronTrack();
        et->ReSet();
        pt = et->GetTrack();
        pt->SetCharge(t.ptype == mine::HEPEM_ELECTRON ? -1.0 : 1.0);
        G4HepEmMSCTrackData* m = et->GetMSCTrackData();
        m->fInitialRange = t.mscInitRange;
        m->fDynamicRangeFactor = t.mscDynRangeFactor;
        m->fTlimitMin = t.mscTlimitMin;
        m->fIsFirstStep = t.mscFirstStep;
        pt->SetEKin(t.ekin);
        pt->SetDirection(t.dir[0], t.dir[1], t.dir[2]);
        pt->SetMCIndex(t.imc);
        pt->SetOnBoundary(t.onBoundary);
        for (int k = 0; k < 4; ++k) pt->SetNumIALeft(t.nial[k], k);
        SlabGeo g;
        g.x = t.x;
        g.lo = lo;
        g.hi = hi;
        g.exit_lo = l == 0;
        g.exit_hi = l == nbins - 1;
        ElectronSlabStep(d, p, &tl, g);
        x = g.x;
        sd = g.side;
        onb = sd != 0;
        left = g.left;
        t.mscInitRange = m->fInitialRange;
        t.mscDynRangeFactor = m->fDynamicRangeFactor;
        t.mscTlimitMin = m->fTlimitMin;
        t.mscFirstStep = m->fIsFirstStep;
      } else {
        G4HepEmGammaTrack* gt = tl.GetPrimaryGammaTrack();
        gt->ReSet();
        pt = gt->GetTrack();
        pt->SetEKin(t.ekin);
        pt->SetDirection(t.dir[0], t.dir[1], t.dir[2]);
        pt->SetMCIndex(t.imc);
        for (int k = 0; k < 4; ++k) pt->SetNumIALeft(t.nial[k], k);
        G4HepEmGammaManager::HowFar(d, p, &tl);
        const double physicalStep = pt->GetGStepLength();
        const double ux = t.dir[0];
        double geometryStep = std::numeric_limits<double>::max();
        if (ux > 0.0) { geometryStep = (hi - t.x) / ux; sd = 1; }
        else if (ux < 0.0) { geometryStep = (lo - t.x) / ux; sd = -1; }
        onb = geometryStep < physicalStep;
        const double finalStep = onb ? geometryStep : physicalStep;
        x = onb ? (sd > 0 ? hi : lo) : std::min(std::max(t.x + finalStep * ux, lo), hi);
        if (!onb) sd = 0;
        pt->SetGStepLength(finalStep);
        pt->SetOnBoundary(onb);
        pt->SetEnergyDeposit(0.0);
        if (onb) {
          G4HepEmGammaManager::UpdateNumIALeft(pt);
        } else {
          G4HepEmGammaManager::SelectInteraction(d, &tl);
          if (pt->GetWinnerProcessIndex() != 3) G4HepEmGammaManager::Perform(d, p, &tl);
        }
      }
      ++steps;
      edep = pt->GetEnergyDeposit();
      const double ekin = pt->GetEKin();
      // the track's secondaries, electron list first (StackSecondaries order)
      const int nel = int(tl.GetNumSecondaryElectronTrack()), ngm = int(tl.GetNumSecondaryGammaTrack());
      struct Sec { int ptype; double ekin, dir[3]; };
      std::vector<Sec> secs;
      for (int k = 0; k < nel + ngm; ++k) {
        G4HepEmTrack* s = k < nel ? tl.GetSecondaryElectronTrack(k)->GetTrack() : tl.GetSecondaryGammaTrack(k - nel)->GetTrack();
        Sec c{k < nel ? (s->GetCharge() < 0.0 ? mine::HEPEM_ELECTRON : mine::HEPEM_POSITRON) : mine::HEPEM_GAMMA, s->GetEKin(),
              {s->GetDirection()[0], s->GetDirection()[1], s->GetDirection()[2]}};
        secs.push_back(c);
      }
      bool alive = ekin > 0.0;
      if (left || (alive && onb && (l + sd < 0 || l + sd >= nbins))) {
        double e = ekin + (t.ptype == mine::HEPEM_POSITRON && ekin > 0.0 ? 2.0 * kMe : 0.0);
        for (const Sec& c : secs) e += c.ekin;
        run.leak[ev] += (int64_t)std::llrint(e * kShFix);
        alive = false;
        secs.clear();
      } else if (alive && onb) {
        l += sd;
      }
      // StackSecondaries: below-threshold secondaries are deposited where the region applies cuts
      const G4HepEmMCCData& mcc = d->fTheMatCutData->fMatCutData[run.su->imc[l0 % nabs]];
      const bool isApplyCuts = p->fParametersPerRegion[mcc.fG4RegionIndex].fIsApplyCuts;
      std::vector<int> stacked;
      for (int s = 0; s < (int)secs.size(); ++s) {
        const Sec& c = secs[s];
        if (isApplyCuts) {
          if (c.ptype == mine::HEPEM_ELECTRON && c.ekin < mcc.fSecElProdCutE) { edep += c.ekin; continue; }
          if (c.ptype == mine::HEPEM_POSITRON && kMe < mcc.fSecGamProdCutE && c.ekin < mcc.fSecPosProdCutE) {
            edep += c.ekin + 2 * kMe;
            continue;
          }
          if (c.ptype == mine::HEPEM_GAMMA && c.ekin < mcc.fSecGamProdCutE) { edep += c.ekin; continue; }
        }
        stacked.push_back(s);
      }
      if (edep > 0.0) dep[l0] += (int64_t)std::llrint(edep * kShFix);
      const int ns1 = t.nstep + 1;
      for (int k = (int)stacked.size() - 1; k >= 0; --k) {
        const int s = stacked[k];
        ShTrack c;
        c.ptype = secs[s].ptype;
        c.stream = ShMix64(t.stream ^ ShMix64(((uint64_t)(uint32_t)ns1 << 2) | (uint64_t)(s + 1)));
        c.imc = run.su->imc[l % nabs];
        c.ekin = secs[s].ekin;
        for (int q = 0; q < 3; ++q) c.dir[q] = secs[s].dir[q];
        c.x = x;
        c.layer = l;
        stack.push_back(c);
        ++tracks;
      }
      if (!alive) break;
      t.ekin = ekin;
      for (int q = 0; q < 3; ++q) t.dir[q] = pt->GetDirection()[q];
      for (int k = 0; k < 4; ++k) t.nial[k] = pt->GetNumIALeft(k);
      t.onBoundary = onb;
      t.ctr = st.ctr;
      t.hasGauss = em->isGauss;
      t.gauss = em->gauss;
      t.x = x;
      t.layer = l;
      t.imc = run.su->imc[l % nabs];
      t.nstep = ns1;
    }
  }
  run.steps += steps;
  run.tracks += tracks;
}
}  // namespace

extern "C" {

int ref_layout_ok() { return 1; }

// Showers in a stack of slabs as hepem_slab_run defines them, every step taken
// by the reference's own code (ShEvent). dep: [nevents][nlayers*nabs] and
// leak: [nevents] in 1e-9 MeV; steps_tracks (optional): {track-steps, tracks}.
// Woodcock tracking is not offered (returns -6).
int ref_slab_shower(const mine::G4HepEmData* dd, const mine::G4HepEmParameters* pp, const mine::hepem_slab_setup* su,
                    int nevents, uint64_t seed, int64_t* dep, int64_t* leak, int64_t* steps_tracks, int nthreads) {
  if (!dd || !pp || !su || !dep || !leak || nevents < 0 || su->nabs < 1 || su->nabs > HEPEM_SLAB_MAX_ABS || su->nlayers < 1)
    return -1;
  if (su->woodcock_min_ekin > 0.0) return -6;
  ShRun run;
  run.d = R(dd);
  run.p = R(pp);
  run.seed = seed;
  run.su = su;
  const int nbins = su->nlayers * su->nabs;
  run.plane.assign((size_t)nbins + 1, 0.0);
  for (int k = 0; k < nbins; ++k) run.plane[k + 1] = run.plane[k] + su->thickness[k % su->nabs];
  std::memset(dep, 0, (size_t)nevents * nbins * sizeof(int64_t));
  std::memset(leak, 0, (size_t)nevents * sizeof(int64_t));
  run.dep = dep;
  run.leak = leak;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > nevents) nthreads = nevents > 0 ? nevents : 1;
  std::atomic<int> next{0};
  std::vector<std::thread> th;
  for (int k = 0; k < nthreads; ++k)
    th.emplace_back([&run, &next, nevents]() {
      for (int e = next++; e < nevents; e = next++) ShEvent(run, e);
    });
  for (auto& h : th) h.join();
  if (steps_tracks) {
    steps_tracks[0] = run.steps;
    steps_tracks[1] = run.tracks;
  }
  return 0;
}

int ref_electron_slab_step(const mine::G4HepEmData* dd, const mine::G4HepEmParameters* pp, uint64_t seed,
                           mine::hepem_tracks* t, mine::hepem_step_out* res, double* x, const double* lo, const double* hi,
                           const uint8_t* exitf, int8_t* side, uint8_t* left, int32_t* nsub, int nthreads) {
  G4HepEmData* d = R(dd);
  G4HepEmParameters* p = R(pp);
  const int n = t->n;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > n) nthreads = n > 0 ? n : 1;
  std::vector<std::thread> th;
  const int chunk = (n + nthreads - 1) / nthreads;
  for (int k = 0; k < nthreads; ++k) {
    const int b = std::min(n, k * chunk), e = std::min(n, (k + 1) * chunk);
    th.emplace_back(ElectronSlabRange, d, p, seed, t, res, x, lo, hi, exitf, side, left, nsub, b, e);
  }
  for (auto& h : th) h.join();
  return 0;
}

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
