// g4hepem_amd.hpp -- C++ host layer over the C-ABI (include/g4hepem_amd.h),
// header only, for callers written against G4HepEm's run-time managers.
//
//   g4hepem_amd::ElectronManager::HowFar / Perform
//       mirror G4HepEmElectronManager::HowFar(G4HepEmData*, G4HepEmParameters*,
//       G4HepEmTLData*) and ::Perform(...) (G4HepEm/G4HepEmRun/include/
//       G4HepEmElectronManager.hh:71,222) for a batch of e-/e+ tracks;
//   g4hepem_amd::GammaManager::HowFar / Perform
//       mirror G4HepEmGammaManager::HowFar and SelectInteraction + Perform
//       (G4HepEmGammaManager.hh:32-55) for a batch of gammas;
//   g4hepem_amd::Context
//       holds the device copy of G4HepEmData + G4HepEmParameters, as
//       CopyG4HepEmDataToGPU does (G4HepEm/G4HepEmData/src/G4HepEmData.cc:78-102).
//
// The reference's managers are static and take the thread-local track
// (G4HepEmTLData); here the state of a whole batch lives in device memory
// (TrackBatch, HowFarState, StepResults) and the calls enqueue kernels on a
// HIP stream. Errors are reported as g4hepem_amd::Error exceptions carrying the
// C-ABI status (the reference reports none); nothing aborts the process.
#ifndef G4HEPEM_AMD_HPP
#define G4HEPEM_AMD_HPP

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "g4hepem_amd.h"

namespace g4hepem_amd {

class Error : public std::runtime_error {
 public:
  Error(const std::string& where, int status)
      : std::runtime_error(where + ": " + hepem_status_string(status)), fStatus(status) {}
  int Status() const { return fStatus; }

 private:
  int fStatus;
};

inline void Check(const char* where, int status) {
  if (status != HEPEM_OK) throw Error(where, status);
}
inline void CheckHip(const char* where, hipError_t e) {
  if (e != hipSuccess) throw Error(std::string(where) + " (" + hipGetErrorString(e) + ")", HEPEM_ERR_HIP);
}

// Device memory owned by one object, freed with it.
class DeviceArena {
 public:
  DeviceArena() = default;
  DeviceArena(const DeviceArena&) = delete;
  DeviceArena& operator=(const DeviceArena&) = delete;
  DeviceArena(DeviceArena&& o) noexcept : fPtrs(std::move(o.fPtrs)) { o.fPtrs.clear(); }
  ~DeviceArena() {
    for (void* p : fPtrs) (void)hipFree(p);
  }
  template <typename T> T* Alloc(size_t n) {
    void* p = nullptr;
    CheckHip("hipMalloc", hipMalloc(&p, (n ? n : 1) * sizeof(T)));
    fPtrs.push_back(p);
    return static_cast<T*>(p);
  }

 private:
  std::vector<void*> fPtrs;
};

enum class Precision { kFP64 = HEPEM_FP64, kFP32Tables = HEPEM_FP32T };

// Device image of the tables (immutable; share it between threads and streams).
class Context {
 public:
  Context(const G4HepEmData* data, const G4HepEmParameters* pars, Precision prec = Precision::kFP64) {
    Check("hepem_ctx_create", hepem_ctx_create(data, pars, static_cast<int>(prec), &fCtx));
  }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  ~Context() {
    if (fCtx) (void)hepem_ctx_destroy(fCtx);
  }
  const hepem_ctx* Get() const { return fCtx; }

 private:
  hepem_ctx* fCtx = nullptr;
};

// Per-track state of a batch in device memory (the SoA of hepem_tracks).
class TrackBatch {
 public:
  explicit TrackBatch(int32_t n, bool withStreamHigh = false) {
    fT.n = n;
    const size_t m = static_cast<size_t>(n);
    fT.ptype = fA.Alloc<int8_t>(m);
    fT.id = fA.Alloc<int32_t>(m);
    fT.imc = fA.Alloc<int32_t>(m);
    fT.ekin = fA.Alloc<double>(m);
    fT.dir = fA.Alloc<double>(3 * m);
    fT.safety = fA.Alloc<double>(m);
    fT.on_boundary = fA.Alloc<uint8_t>(m);
    fT.num_ia_left = fA.Alloc<double>(4 * m);
    fT.msc_init_range = fA.Alloc<double>(m);
    fT.msc_dyn_range_factor = fA.Alloc<double>(m);
    fT.msc_tlimit_min = fA.Alloc<double>(m);
    fT.msc_first_step = fA.Alloc<uint8_t>(m);
    fT.rng_ctr = fA.Alloc<uint32_t>(m);
    fT.rng_gauss = fA.Alloc<double>(m);
    fT.rng_has_gauss = fA.Alloc<uint8_t>(m);
    fT.rng_stream = withStreamHigh ? fA.Alloc<uint32_t>(m) : nullptr;
  }
  int32_t Size() const { return fT.n; }
  hepem_tracks& View() { return fT; }
  const hepem_tracks& View() const { return fT; }
  // G4HepEmTrack::ReSet + G4HepEmMSCTrackData::ReSet + DiscardGauss of every track
  void ResetPhysicsState(hipStream_t s = nullptr) { Check("hepem_tracks_reset_state", hepem_tracks_reset_state(&fT, s)); }

 private:
  DeviceArena fA;
  hepem_tracks fT{};
};

// HowFar -> Perform hand-over of a batch (hepem_howfar_state).
class HowFarState {
 public:
  explicit HowFarState(int32_t n) {
    const size_t m = static_cast<size_t>(n);
    fH.n = n;
    fH.geom_step = fA.Alloc<double>(m);
    fH.range = fA.Alloc<double>(m);
    fH.mfp = fA.Alloc<double>(4 * m);
    fH.winner = fA.Alloc<int32_t>(m);
    fH.msc_lambtr1 = fA.Alloc<double>(m);
    fH.msc_true_step = fA.Alloc<double>(m);
    fH.msc_zpath = fA.Alloc<double>(m);
    fH.msc_par1 = fA.Alloc<double>(m);
    fH.msc_par2 = fA.Alloc<double>(m);
    fH.msc_par3 = fA.Alloc<double>(m);
    fH.msc_active = fA.Alloc<uint8_t>(m);
    fH.msc_displace = fA.Alloc<uint8_t>(m);
    fH.pe_xsec = fA.Alloc<double>(m);
  }
  hepem_howfar_state& View() { return fH; }
  // proposed geometrical step of each track (device array): the caller's
  // geometry may shorten it before Perform, as G4HepEmTrackingManager does
  double* GeomStep() { return fH.geom_step; }

 private:
  DeviceArena fA;
  hepem_howfar_state fH{};
};

// Per-step results of a batch (hepem_step_out), with up to two secondaries
// per track in fixed slots.
class StepResults {
 public:
  explicit StepResults(int32_t n) {
    const size_t m = static_cast<size_t>(n);
    fR.true_step = fA.Alloc<double>(m);
    fR.geom_step = fA.Alloc<double>(m);
    fR.edep = fA.Alloc<double>(m);
    fR.displacement = fA.Alloc<double>(3 * m);
    fR.winner = fA.Alloc<int32_t>(m);
    fR.stopped = fA.Alloc<uint8_t>(m);
    fR.nsec = fA.Alloc<int32_t>(m);
    fR.sec_ptype = fA.Alloc<int8_t>(2 * m);
    fR.sec_ekin = fA.Alloc<double>(2 * m);
    fR.sec_dir = fA.Alloc<double>(6 * m);
  }
  hepem_step_out& View() { return fR; }

 private:
  DeviceArena fA;
  hepem_step_out fR{};
};

// The particle-type queue and counters of a batch's steps.
class WorkBuffer {
 public:
  explicit WorkBuffer(int32_t n) : fPtr(fA.Alloc<uint8_t>(hepem_step_work_bytes(n))) {}
  void* Get() { return fPtr; }
  // tracks the last HowFar/step on this buffer rejected (synchronises s)
  uint32_t Rejected(hipStream_t s = nullptr) const {
    uint32_t n = 0;
    Check("hepem_step_rejected", hepem_step_rejected(fPtr, &n, s));
    return n;
  }

 private:
  DeviceArena fA;
  void* fPtr;
};

// G4HepEmElectronManager for a batch of e-/e+ (every slot must hold one).
struct ElectronManager {
  static void HowFar(const Context& ctx, uint64_t seed, TrackBatch& t, HowFarState& hs, WorkBuffer& w,
                     hipStream_t s = nullptr) {
    Check("hepem_electron_howfar", hepem_electron_howfar(ctx.Get(), seed, &t.View(), &hs.View(), w.Get(), s));
  }
  static void Perform(const Context& ctx, uint64_t seed, TrackBatch& t, HowFarState& hs, StepResults& r, WorkBuffer& w,
                      hipStream_t s = nullptr) {
    Check("hepem_electron_perform",
          hepem_electron_perform(ctx.Get(), seed, &t.View(), &hs.View(), &r.View(), w.Get(), s));
  }
};

// G4HepEmGammaManager for a batch of gammas (every slot must hold one).
struct GammaManager {
  static void HowFar(const Context& ctx, uint64_t seed, TrackBatch& t, HowFarState& hs, WorkBuffer& w,
                     hipStream_t s = nullptr) {
    Check("hepem_gamma_howfar", hepem_gamma_howfar(ctx.Get(), seed, &t.View(), &hs.View(), w.Get(), s));
  }
  // SelectInteraction + Perform, or only the interaction-length update for
  // tracks whose post-step point is on a boundary
  static void Perform(const Context& ctx, uint64_t seed, TrackBatch& t, HowFarState& hs, StepResults& r, WorkBuffer& w,
                      hipStream_t s = nullptr) {
    Check("hepem_gamma_perform", hepem_gamma_perform(ctx.Get(), seed, &t.View(), &hs.View(), &r.View(), w.Get(), s));
  }
};

// Both managers over a mixed batch: HowFar / Perform of every track, or the
// whole step with no geometry in between.
struct Stepper {
  static void HowFar(const Context& ctx, uint64_t seed, TrackBatch& t, HowFarState& hs, WorkBuffer& w,
                     hipStream_t s = nullptr) {
    Check("hepem_howfar", hepem_howfar(ctx.Get(), seed, &t.View(), &hs.View(), w.Get(), s));
  }
  static void Perform(const Context& ctx, uint64_t seed, TrackBatch& t, HowFarState& hs, StepResults& r, WorkBuffer& w,
                      hipStream_t s = nullptr) {
    Check("hepem_perform", hepem_perform(ctx.Get(), seed, &t.View(), &hs.View(), &r.View(), w.Get(), s));
  }
  static void Step(const Context& ctx, uint64_t seed, const TrackBatch& in, TrackBatch& out, StepResults& r, WorkBuffer& w,
                   hipStream_t s = nullptr) {
    Check("hepem_step", hepem_step(ctx.Get(), seed, &in.View(), &out.View(), &r.View(), w.Get(), s));
  }
};

}  // namespace g4hepem_amd

#endif  // G4HEPEM_AMD_HPP
