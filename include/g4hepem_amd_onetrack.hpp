// g4hepem_amd_onetrack.hpp -- one-track drop-in for G4HepEm's run-time
// managers, header only, over the C-ABI (include/g4hepem_amd.h).
//
// The classes a TrackingManager-style caller holds, with the reference's
// names, methods and TLData semantics (G4HepEm/G4HepEmRun/include/):
//   G4HepEmTrack, G4HepEmMSCTrackData        (G4HepEmTrack.hh, G4HepEmMSCTrackData.hh)
//   G4HepEmElectronTrack, G4HepEmGammaTrack  (G4HepEmElectronTrack.hh, G4HepEmGammaTrack.hh)
//   G4HepEmTLData                            (G4HepEmTLData.hh)
//   G4HepEmElectronManager::HowFar / Perform (G4HepEmElectronManager.hh:71,222)
//   G4HepEmGammaManager::HowFar / Perform    (G4HepEmGammaManager.hh:32,39)
// Each manager call steps the thread's one primary track on the GPU through
// the C-ABI (hepem_electron_howfar / _perform, hepem_gamma_howfar / _perform)
// and returns when its results are back in the TLData, so a caller written
// against the reference works unchanged, one track at a time. That is for
// porting and checking: throughput needs batches (g4hepem_amd.hpp or the
// C-ABI). Differences from the reference:
//   * the random numbers come from the per-track Philox stream of the C-ABI
//     (G4HepEmTLData::SetSeed and the track's ID select it; its counter and
//     Gaussian cache travel with the track), not from a G4HepEmRandomEngine;
//   * the G4HepEmData / G4HepEmParameters pointers a manager is called with
//     select a device image of those tables, made at the first call with that
//     pair and kept until the process ends (the tables must not change);
//   * errors raise g4hepem_amd::Error (g4hepem_amd.hpp) with the C-ABI status;
//   * G4HepEmGammaManager::Perform takes the boundary update itself when the
//     track is on a boundary (the reference caller calls UpdateNumIALeft).
#ifndef G4HEPEM_AMD_ONETRACK_HPP
#define G4HEPEM_AMD_ONETRACK_HPP

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "g4hepem_amd.hpp"

namespace g4hepem_amd {
namespace onetrack {

// G4HepEmTrack (G4HepEmTrack.hh): the neutral part of a track, plus the state
// of its Philox stream on the device (RngCounter, RngGauss, RngHasGauss,
// RngStream), which the reference keeps in its random engine instead.
class G4HepEmTrack {
 public:
  G4HepEmTrack() { ReSet(); }
  void SetPosition(const double* p) { SetPosition(p[0], p[1], p[2]); }
  void SetPosition(double x, double y, double z) {
    fPosition[0] = x;
    fPosition[1] = y;
    fPosition[2] = z;
  }
  double* GetPosition() { return fPosition; }
  void SetDirection(const double* d) { SetDirection(d[0], d[1], d[2]); }
  void SetDirection(double x, double y, double z) {
    fDirection[0] = x;
    fDirection[1] = y;
    fDirection[2] = z;
  }
  double* GetDirection() { return fDirection; }
  void SetEKin(double ekin) {
    fEKin = ekin;
    fLogEKin = 100.0;  // recomputed when asked for
  }
  void SetLEKin(double lekin) { fLogEKin = lekin; }
  void SetEKin(double ekin, double lekin) {
    fEKin = ekin;
    fLogEKin = lekin;
  }
  double GetEKin() const { return fEKin; }
  double GetLogEKin() {
    if (fLogEKin > 99.0) fLogEKin = fEKin > 0.0 ? std::log(fEKin) : -30.0;
    return fLogEKin;
  }
  void SetCharge(double ch) { fCharge = ch; }
  double GetCharge() const { return fCharge; }
  void SetEnergyDeposit(double v) { fEDeposit = v; }
  void AddEnergyDeposit(double v) { fEDeposit += v; }
  double GetEnergyDeposit() const { return fEDeposit; }
  void SetGStepLength(double v) { fGStepLength = v; }
  double GetGStepLength() const { return fGStepLength; }
  void SetMFP(double v, int k) { fMFPs[k] = v; }
  double GetMFP(int k) const { return fMFPs[k]; }
  double* GetMFP() { return fMFPs; }
  void SetNumIALeft(double v, int k) { fNumIALeft[k] = v; }
  double GetNumIALeft(int k) const { return fNumIALeft[k]; }
  double* GetNumIALeft() { return fNumIALeft; }
  void SetSafety(double s) { fSafety = s; }
  double GetSafety() const { return fSafety; }
  void SetID(int id) { fID = id; }
  int GetID() const { return fID; }
  void SetParentID(int id) { fIDParent = id; }
  int GetParentID() const { return fIDParent; }
  void SetMCIndex(int imc) { fMCIndex = imc; }
  int GetMCIndex() const { return fMCIndex; }
  void SetWinnerProcessIndex(int ip) { fWinner = ip; }
  int GetWinnerProcessIndex() const { return fWinner; }
  void SetOnBoundary(bool v) { fOnBoundary = v; }
  bool GetOnBoundary() const { return fOnBoundary; }
  // the track's Philox stream: high word of the stream selector, draws taken,
  // cached second Gaussian deviate
  void SetRngStream(uint32_t s) { fRngStream = s; }
  uint32_t GetRngStream() const { return fRngStream; }
  uint32_t GetRngCounter() const { return fRngCtr; }
  void SetRngState(uint32_t ctr, double gauss, bool hasGauss) {
    fRngCtr = ctr;
    fRngGauss = gauss;
    fRngHasGauss = hasGauss;
  }
  double GetRngGauss() const { return fRngGauss; }
  bool GetRngHasGauss() const { return fRngHasGauss; }
  // start-of-track values (G4HepEmTrack::ReSet), the stream at its start
  void ReSet() {
    for (int c = 0; c < 3; ++c) fPosition[c] = fDirection[c] = 0.0;
    fEKin = 0.0;
    fLogEKin = 100.0;
    fCharge = 0.0;
    fEDeposit = 0.0;
    fGStepLength = 0.0;
    for (int k = 0; k < 4; ++k) fMFPs[k] = fNumIALeft[k] = -1.0;
    fSafety = 0.0;
    fID = fIDParent = fMCIndex = fWinner = -1;
    fOnBoundary = false;
    fRngStream = 0u;
    fRngCtr = 0u;
    fRngGauss = 0.0;
    fRngHasGauss = false;
  }

 private:
  double fPosition[3], fDirection[3];
  double fEKin, fLogEKin, fCharge, fEDeposit, fGStepLength;
  double fMFPs[4], fNumIALeft[4];
  double fSafety;
  int fID, fIDParent, fMCIndex, fWinner;
  bool fOnBoundary;
  uint32_t fRngStream, fRngCtr;
  double fRngGauss;
  bool fRngHasGauss;
};

// G4HepEmMSCTrackData (G4HepEmMSCTrackData.hh): public data members, as there.
class G4HepEmMSCTrackData {
 public:
  G4HepEmMSCTrackData() { ReSet(); }
  void SetDisplacement(double x, double y, double z) {
    fDisplacement[0] = x;
    fDisplacement[1] = y;
    fDisplacement[2] = z;
  }
  double* GetDisplacement() { return fDisplacement; }
  void SetNewDirection(double x, double y, double z) {
    fDirection[0] = x;
    fDirection[1] = y;
    fDirection[2] = z;
  }
  double* GetDirection() { return fDirection; }
  void ReSet() {
    fLambtr1 = fTrueStepLength = fZPathLength = 0.0;
    fDisplacement[0] = fDisplacement[1] = fDisplacement[2] = 0.0;
    fDirection[0] = fDirection[1] = 0.0;
    fDirection[2] = 1.0;
    fInitialRange = 1.0e+21;
    fDynamicRangeFactor = 0.04;
    fTlimitMin = 1.0E-7;
    fPar1 = -1.0;
    fPar2 = fPar3 = 0.0;
    fIsNoScatteringInMSC = fIsDisplace = fIsActive = false;
    fIsFirstStep = true;
  }

  double fLambtr1, fTrueStepLength, fZPathLength;
  double fDisplacement[3], fDirection[3];
  double fInitialRange, fDynamicRangeFactor, fTlimitMin;
  double fPar1, fPar2, fPar3;
  bool fIsNoScatteringInMSC, fIsDisplace, fIsFirstStep, fIsActive;
};

// G4HepEmElectronTrack (G4HepEmElectronTrack.hh)
class G4HepEmElectronTrack {
 public:
  G4HepEmElectronTrack() { ReSet(); }
  G4HepEmTrack* GetTrack() { return &fTrack; }
  G4HepEmMSCTrackData* GetMSCTrackData() { return &fMSC; }
  double GetCharge() const { return fTrack.GetCharge(); }
  void SetRange(double r) { fRange = r; }
  double GetRange() const { return fRange; }
  void SetPStepLength(double s) { fPStepLength = s; }
  double GetPStepLength() const { return fPStepLength; }
  void SavePreStepEKin() {
    fPreStepEKin = fTrack.GetEKin();
    fPreStepLogEKin = fTrack.GetLogEKin();
  }
  void SetPreStepEKin(double ekin, double lekin) {
    fPreStepEKin = ekin;
    fPreStepLogEKin = lekin;
  }
  double GetPreStepEKin() const { return fPreStepEKin; }
  double GetPreStepLogEKin() const { return fPreStepLogEKin; }
  void ReSet() {
    fTrack.ReSet();
    fTrack.SetCharge(-1.0);
    fMSC.ReSet();
    fRange = fPStepLength = 0.0;
    fPreStepEKin = 0.0;
    fPreStepLogEKin = 100.0;
  }

 private:
  G4HepEmTrack fTrack;
  G4HepEmMSCTrackData fMSC;
  double fRange, fPStepLength, fPreStepEKin, fPreStepLogEKin;
};

// G4HepEmGammaTrack (G4HepEmGammaTrack.hh)
class G4HepEmGammaTrack {
 public:
  G4HepEmGammaTrack() { ReSet(); }
  G4HepEmTrack* GetTrack() { return &fTrack; }
  void SetPEmxSec(double v) { fPEmxSec = v; }
  double GetPEmxSec() const { return fPEmxSec; }
  void ReSet() {
    fTrack.ReSet();
    fPEmxSec = 0.0;
  }

 private:
  G4HepEmTrack fTrack;
  double fPEmxSec;
};

namespace detail {
// One track's C-ABI arrays in one block of device memory, mirrored on the
// host: hepem_tracks, hepem_howfar_state and hepem_step_out with n = 1 (a
// planar field of k components is k consecutive values).
struct OneTrack {
  double ekin, dir[3], safety, nial[4], mir, mdrf, mtm, gauss;
  double hGeom, hRange, hMfp[4], hLambtr1, hTrue, hZpath, hPar1, hPar2, hPar3, hPe;
  double rTrue, rGeom, rEdep, rDispl[3], rSecEkin[2], rSecDir[6];
  int32_t id, imc, hWinner, rWinner, rNsec;
  uint32_t ctr, stream;
  int8_t ptype, rSecPtype[2];
  uint8_t onb, mfs, hasGauss, hActive, hDisplace, rStopped;
};

// the views of a OneTrack block at `b`
inline void Views(OneTrack* b, hepem_tracks& t, hepem_howfar_state& h, hepem_step_out& r) {
  t = hepem_tracks{};
  t.n = 1;
  t.ptype = &b->ptype;
  t.id = &b->id;
  t.imc = &b->imc;
  t.ekin = &b->ekin;
  t.dir = b->dir;
  t.safety = &b->safety;
  t.on_boundary = &b->onb;
  t.num_ia_left = b->nial;
  t.msc_init_range = &b->mir;
  t.msc_dyn_range_factor = &b->mdrf;
  t.msc_tlimit_min = &b->mtm;
  t.msc_first_step = &b->mfs;
  t.rng_ctr = &b->ctr;
  t.rng_gauss = &b->gauss;
  t.rng_has_gauss = &b->hasGauss;
  t.rng_stream = &b->stream;
  h = hepem_howfar_state{};
  h.n = 1;
  h.geom_step = &b->hGeom;
  h.range = &b->hRange;
  h.mfp = b->hMfp;
  h.winner = &b->hWinner;
  h.msc_lambtr1 = &b->hLambtr1;
  h.msc_true_step = &b->hTrue;
  h.msc_zpath = &b->hZpath;
  h.msc_par1 = &b->hPar1;
  h.msc_par2 = &b->hPar2;
  h.msc_par3 = &b->hPar3;
  h.msc_active = &b->hActive;
  h.msc_displace = &b->hDisplace;
  h.pe_xsec = &b->hPe;
  r = hepem_step_out{};
  r.true_step = &b->rTrue;
  r.geom_step = &b->rGeom;
  r.edep = &b->rEdep;
  r.displacement = b->rDispl;
  r.winner = &b->rWinner;
  r.stopped = &b->rStopped;
  r.nsec = &b->rNsec;
  r.sec_ptype = b->rSecPtype;
  r.sec_ekin = b->rSecEkin;
  r.sec_dir = b->rSecDir;
}

// Device images of the (G4HepEmData, G4HepEmParameters) pairs the managers
// were called with, made at first use and kept until the process ends.
// (never freed: they outlive every static object that might still step)
inline const hepem_ctx* ContextFor(const G4HepEmData* data, const G4HepEmParameters* pars) {
  static std::mutex m;
  static auto* images = new std::map<std::pair<const void*, const void*>, Context*>();
  std::lock_guard<std::mutex> lock(m);
  Context*& c = (*images)[{data, pars}];
  if (!c) c = new Context(data, pars);
  return c->Get();
}
}  // namespace detail

// G4HepEmTLData (G4HepEmTLData.hh): the thread's primary tracks and the
// secondaries its steps add, which the caller drains and resets; plus the
// seed of the Philox streams and the device block of the one-track calls.
class G4HepEmTLData {
 public:
  G4HepEmTLData() : fElectronSecondaries(2), fGammaSecondaries(2) {}
  G4HepEmTLData(const G4HepEmTLData&) = delete;
  G4HepEmTLData& operator=(const G4HepEmTLData&) = delete;
  ~G4HepEmTLData() {
    if (fDev) (void)hipFree(fDev);
    if (fWork) (void)hipFree(fWork);
  }
  void SetSeed(uint64_t seed) { fSeed = seed; }
  uint64_t GetSeed() const { return fSeed; }

  G4HepEmElectronTrack* GetPrimaryElectronTrack() { return &fElectron; }
  G4HepEmElectronTrack* AddSecondaryElectronTrack() {
    if (fNumElectronSecondaries == fElectronSecondaries.size()) fElectronSecondaries.resize(2 * fElectronSecondaries.size());
    return &fElectronSecondaries[fNumElectronSecondaries++];
  }
  std::size_t GetNumSecondaryElectronTrack() const { return fNumElectronSecondaries; }
  void ResetNumSecondaryElectronTrack() { fNumElectronSecondaries = 0; }
  G4HepEmElectronTrack* GetSecondaryElectronTrack(int k) { return &fElectronSecondaries[(size_t)k]; }

  G4HepEmGammaTrack* GetPrimaryGammaTrack() { return &fGamma; }
  G4HepEmGammaTrack* AddSecondaryGammaTrack() {
    if (fNumGammaSecondaries == fGammaSecondaries.size()) fGammaSecondaries.resize(2 * fGammaSecondaries.size());
    return &fGammaSecondaries[fNumGammaSecondaries++];
  }
  std::size_t GetNumSecondaryGammaTrack() const { return fNumGammaSecondaries; }
  void ResetNumSecondaryGammaTrack() { fNumGammaSecondaries = 0; }
  G4HepEmGammaTrack* GetSecondaryGammaTrack(int k) { return &fGammaSecondaries[(size_t)k]; }

  // the one-track block: uploaded, `call` made on its views, downloaded
  template <typename F> void Call(const char* where, F&& call) {
    if (!fDev) {
      CheckHip("hipMalloc", hipMalloc(reinterpret_cast<void**>(&fDev), sizeof(detail::OneTrack)));
      CheckHip("hipMalloc", hipMalloc(&fWork, hepem_step_work_bytes(1)));
    }
    CheckHip("hipMemcpy", hipMemcpy(fDev, &fHost, sizeof(fHost), hipMemcpyHostToDevice));
    hepem_tracks t;
    hepem_howfar_state h;
    hepem_step_out r;
    detail::Views(fDev, t, h, r);
    Check(where, call(t, h, r, fWork));
    CheckHip("hipMemcpy", hipMemcpy(&fHost, fDev, sizeof(fHost), hipMemcpyDeviceToHost));
  }
  detail::OneTrack& Host() { return fHost; }

 private:
  G4HepEmElectronTrack fElectron;
  G4HepEmGammaTrack fGamma;
  std::vector<G4HepEmElectronTrack> fElectronSecondaries;
  std::vector<G4HepEmGammaTrack> fGammaSecondaries;
  std::size_t fNumElectronSecondaries = 0, fNumGammaSecondaries = 0;
  uint64_t fSeed = 0;
  detail::OneTrack fHost{};
  detail::OneTrack* fDev = nullptr;
  void* fWork = nullptr;
};

namespace detail {
// the persistent state of track t (+ MSC data) into the block, and back
inline void PutTrack(OneTrack& b, G4HepEmTrack& t, int8_t ptype, const G4HepEmMSCTrackData* msc) {
  b.ptype = ptype;
  b.id = t.GetID();
  b.stream = t.GetRngStream();
  b.imc = t.GetMCIndex();
  b.ekin = t.GetEKin();
  for (int c = 0; c < 3; ++c) b.dir[c] = t.GetDirection()[c];
  b.safety = t.GetSafety();
  b.onb = t.GetOnBoundary() ? 1 : 0;
  for (int k = 0; k < 4; ++k) b.nial[k] = t.GetNumIALeft(k);
  b.ctr = t.GetRngCounter();
  b.gauss = t.GetRngGauss();
  b.hasGauss = t.GetRngHasGauss() ? 1 : 0;
  G4HepEmMSCTrackData fresh;
  const G4HepEmMSCTrackData& m = msc ? *msc : fresh;
  b.mir = m.fInitialRange;
  b.mdrf = m.fDynamicRangeFactor;
  b.mtm = m.fTlimitMin;
  b.mfs = m.fIsFirstStep ? 1 : 0;
}
inline void GetTrack(const OneTrack& b, G4HepEmTrack& t, G4HepEmMSCTrackData* msc) {
  if (b.ekin != t.GetEKin()) t.SetEKin(b.ekin);
  t.SetDirection(b.dir[0], b.dir[1], b.dir[2]);
  for (int k = 0; k < 4; ++k) t.SetNumIALeft(b.nial[k], k);
  t.SetRngState(b.ctr, b.gauss, b.hasGauss != 0);
  if (msc) {
    msc->fInitialRange = b.mir;
    msc->fDynamicRangeFactor = b.mdrf;
    msc->fTlimitMin = b.mtm;
    msc->fIsFirstStep = b.mfs != 0;
  }
}
// the step's secondaries into the TLData's lists, electrons first, as the
// reference's interactions add them
inline void GetSecondaries(const OneTrack& b, G4HepEmTLData* tl, int parentID) {
  for (int s = 0; s < b.rNsec && s < 2; ++s) {
    G4HepEmTrack* st;
    if (b.rSecPtype[s] == HEPEM_GAMMA) {
      G4HepEmGammaTrack* g = tl->AddSecondaryGammaTrack();
      g->ReSet();
      st = g->GetTrack();
      st->SetCharge(0.0);
    } else {
      G4HepEmElectronTrack* e = tl->AddSecondaryElectronTrack();
      e->ReSet();
      st = e->GetTrack();
      st->SetCharge(b.rSecPtype[s] == HEPEM_ELECTRON ? -1.0 : 1.0);
    }
    st->SetEKin(b.rSecEkin[s]);
    st->SetDirection(b.rSecDir[s], b.rSecDir[2 + s], b.rSecDir[4 + s]);
    st->SetParentID(parentID);
  }
}
}  // namespace detail

// G4HepEmElectronManager (G4HepEmElectronManager.hh:71,222) for the TLData's
// primary e-/e+ (charge -1 / +1).
struct G4HepEmElectronManager {
  // the physics step limit: the track's geometrical step, MFPs and winner,
  // the range and physical step, the MSC step data
  static void HowFar(G4HepEmData* data, G4HepEmParameters* pars, G4HepEmTLData* tl) {
    const hepem_ctx* ctx = detail::ContextFor(data, pars);
    G4HepEmElectronTrack* et = tl->GetPrimaryElectronTrack();
    G4HepEmTrack& t = *et->GetTrack();
    G4HepEmMSCTrackData& m = *et->GetMSCTrackData();
    detail::OneTrack& b = tl->Host();
    detail::PutTrack(b, t, t.GetCharge() < 0.0 ? HEPEM_ELECTRON : HEPEM_POSITRON, &m);
    const uint64_t seed = tl->GetSeed();
    tl->Call("hepem_electron_howfar", [&](hepem_tracks& dt, hepem_howfar_state& dh, hepem_step_out&, void* w) {
      return hepem_electron_howfar(ctx, seed, &dt, &dh, w, nullptr);
    });
    detail::GetTrack(b, t, &m);
    t.SetGStepLength(b.hGeom);
    for (int k = 0; k < 4; ++k) t.SetMFP(b.hMfp[k], k);
    t.SetWinnerProcessIndex(b.hWinner);
    et->SetRange(b.hRange);
    et->SetPStepLength(b.hActive ? b.hTrue : b.hGeom);
    m.fLambtr1 = b.hLambtr1;
    m.fTrueStepLength = b.hTrue;
    m.fZPathLength = b.hZpath;
    m.fPar1 = b.hPar1;
    m.fPar2 = b.hPar2;
    m.fPar3 = b.hPar3;
    m.fIsActive = b.hActive != 0;
    m.fIsDisplace = b.hDisplace != 0;
  }
  // the step over the (possibly geometry-limited) geometrical step length:
  // energy loss, MSC and the discrete interaction; the post-step state, the
  // deposit and the MSC displacement on the track, secondaries in the TLData
  static void Perform(G4HepEmData* data, G4HepEmParameters* pars, G4HepEmTLData* tl) {
    const hepem_ctx* ctx = detail::ContextFor(data, pars);
    G4HepEmElectronTrack* et = tl->GetPrimaryElectronTrack();
    G4HepEmTrack& t = *et->GetTrack();
    G4HepEmMSCTrackData& m = *et->GetMSCTrackData();
    detail::OneTrack& b = tl->Host();
    detail::PutTrack(b, t, t.GetCharge() < 0.0 ? HEPEM_ELECTRON : HEPEM_POSITRON, &m);
    b.hGeom = t.GetGStepLength();
    b.hRange = et->GetRange();
    for (int k = 0; k < 4; ++k) b.hMfp[k] = t.GetMFP(k);
    b.hWinner = t.GetWinnerProcessIndex();
    b.hLambtr1 = m.fLambtr1;
    b.hTrue = m.fTrueStepLength;
    b.hZpath = m.fZPathLength;
    b.hPar1 = m.fPar1;
    b.hPar2 = m.fPar2;
    b.hPar3 = m.fPar3;
    b.hActive = m.fIsActive ? 1 : 0;
    b.hDisplace = m.fIsDisplace ? 1 : 0;
    const uint64_t seed = tl->GetSeed();
    tl->Call("hepem_electron_perform", [&](hepem_tracks& dt, hepem_howfar_state& dh, hepem_step_out& dr, void* w) {
      return hepem_electron_perform(ctx, seed, &dt, &dh, &dr, w, nullptr);
    });
    detail::GetTrack(b, t, &m);
    t.SetEnergyDeposit(b.rEdep);
    t.SetWinnerProcessIndex(b.rWinner);
    et->SetPStepLength(b.rTrue);
    m.SetDisplacement(b.rDispl[0], b.rDispl[1], b.rDispl[2]);
    detail::GetSecondaries(b, tl, t.GetID());
  }
};

// G4HepEmGammaManager (G4HepEmGammaManager.hh:32,39) for the TLData's
// primary gamma.
struct G4HepEmGammaManager {
  // the physics step limit: the track's geometrical step, total MFP, winner
  // and the photoelectric cross section
  static void HowFar(G4HepEmData* data, G4HepEmParameters* pars, G4HepEmTLData* tl) {
    const hepem_ctx* ctx = detail::ContextFor(data, pars);
    G4HepEmGammaTrack* gt = tl->GetPrimaryGammaTrack();
    G4HepEmTrack& t = *gt->GetTrack();
    detail::OneTrack& b = tl->Host();
    detail::PutTrack(b, t, HEPEM_GAMMA, nullptr);
    const uint64_t seed = tl->GetSeed();
    tl->Call("hepem_gamma_howfar", [&](hepem_tracks& dt, hepem_howfar_state& dh, hepem_step_out&, void* w) {
      return hepem_gamma_howfar(ctx, seed, &dt, &dh, w, nullptr);
    });
    detail::GetTrack(b, t, nullptr);
    t.SetGStepLength(b.hGeom);
    t.SetMFP(b.hMfp[0], 0);  // the total (a gamma's only MFP slot)
    t.SetWinnerProcessIndex(b.hWinner);
    gt->SetPEmxSec(b.hPe);
  }
  // SelectInteraction + Perform, or the interaction-length update alone when
  // the track is on a boundary after the step
  static void Perform(G4HepEmData* data, G4HepEmParameters* pars, G4HepEmTLData* tl) {
    const hepem_ctx* ctx = detail::ContextFor(data, pars);
    G4HepEmGammaTrack* gt = tl->GetPrimaryGammaTrack();
    G4HepEmTrack& t = *gt->GetTrack();
    detail::OneTrack& b = tl->Host();
    detail::PutTrack(b, t, HEPEM_GAMMA, nullptr);
    b.hGeom = t.GetGStepLength();
    for (int k = 0; k < 4; ++k) b.hMfp[k] = t.GetMFP(k);
    b.hWinner = t.GetWinnerProcessIndex();
    b.hPe = gt->GetPEmxSec();
    const uint64_t seed = tl->GetSeed();
    tl->Call("hepem_gamma_perform", [&](hepem_tracks& dt, hepem_howfar_state& dh, hepem_step_out& dr, void* w) {
      return hepem_gamma_perform(ctx, seed, &dt, &dh, &dr, w, nullptr);
    });
    detail::GetTrack(b, t, nullptr);
    t.SetEnergyDeposit(b.rEdep);
    t.SetWinnerProcessIndex(b.rWinner);
    detail::GetSecondaries(b, tl, t.GetID());
  }
};

}  // namespace onetrack
}  // namespace g4hepem_amd

#endif  // G4HEPEM_AMD_ONETRACK_HPP
