// dev_soa.hh -- loads and stores of one track slot of the structure-of-arrays
// batches of the C-ABI (include/g4hepem_amd.h), and the whole-step device
// functions of one slot built on them: G4HepEmElectronManager::HowFar + Perform
// and G4HepEmGammaManager::HowFar + SelectInteraction + Perform.
#pragma once

#include "dev_physics.hh"
#include "g4hepem_amd.h"

namespace hepem {

// Secondary slots beyond nsec[i] are left unwritten (the reference's
// G4HepEmTLData likewise holds only the secondaries created).
template <typename R> HD void WriteSlot(const hepem_step_out& res, int n, int i, int s, const Sec<R>& sec) {
  res.sec_ptype[s * n + i] = (int8_t)sec.ptype;
  res.sec_ekin[s * n + i] = sec.ekin;
#pragma unroll
  for (int c = 0; c < 3; ++c) res.sec_dir[(c * 2 + s) * n + i] = (double)sec.dir[c];
}
template <typename R>
HD void WriteSecondaries(const hepem_step_out& res, int n, int i, const SecList<R>& sl) {
  res.nsec[i] = sl.ns;
  if (sl.ns > 0) WriteSlot<R>(res, n, i, 0, sl.s0);
  if (sl.ns > 1) WriteSlot<R>(res, n, i, 1, sl.s1);
}

// A track is stepped only with a valid couple index and a finite, positive
// kinetic energy; otherwise it is passed through unchanged, flagged with
// winner -99 and counted in counters[2].
template <typename R> HD bool TrackValid(const DevTables<R>& T, const hepem_tracks& in, int i) {
  const int imc = in.imc[i];
  const double e = in.ekin[i];
  return imc >= 0 && imc < T.nmc && e > 0.0 && e < 1.0e300;
}

template <typename R>
HD void WriteRejected(const hepem_tracks& in, hepem_tracks& out, const hepem_step_out& res, int n, int i) {
  res.true_step[i] = 0.0;
  res.geom_step[i] = 0.0;
  res.edep[i] = 0.0;
  for (int c = 0; c < 3; ++c) res.displacement[c * n + i] = 0.0;
  res.winner[i] = -99;
  res.stopped[i] = 1;
  SecList<R> sl;
  sl.ns = 0;
  WriteSecondaries(res, n, i, sl);
  if (out.ekin != in.ekin) {
    out.ptype[i] = in.ptype[i];
    out.id[i] = in.id[i];
    if (out.rng_stream) out.rng_stream[i] = in.rng_stream ? in.rng_stream[i] : 0u;
    out.imc[i] = in.imc[i];
    out.ekin[i] = in.ekin[i];
    for (int c = 0; c < 3; ++c) out.dir[c * n + i] = in.dir[c * n + i];
    out.safety[i] = in.safety[i];
    out.on_boundary[i] = in.on_boundary[i];
    for (int k = 0; k < 4; ++k) out.num_ia_left[k * n + i] = in.num_ia_left[k * n + i];
    out.rng_ctr[i] = in.rng_ctr[i];
    out.rng_gauss[i] = in.rng_gauss[i];
    out.rng_has_gauss[i] = in.rng_has_gauss[i];
    if (in.ptype[i] != 2) {
      out.msc_init_range[i] = in.msc_init_range[i];
      out.msc_dyn_range_factor[i] = in.msc_dyn_range_factor[i];
      out.msc_tlimit_min[i] = in.msc_tlimit_min[i];
      out.msc_first_step[i] = in.msc_first_step[i];
    }
  }
}

template <typename R>
HD bool RejectTrack(const DevTables<R>& T, const hepem_tracks& in, hepem_tracks& out, const hepem_step_out& res, int n, int i,
                    uint32_t* counters) {
  if (TrackValid(T, in, i)) return false;
  atomicAdd(&counters[2], 1u);
  WriteRejected<R>(in, out, res, n, i);
  return true;
}

// High word of the stream selector of slot i (0 without the optional array).
HD uint32_t StreamHi(const hepem_tracks& t, int i) { return t.rng_stream ? t.rng_stream[i] : 0u; }

// Copies the identity fields of slot i to a separate output batch.
HD void CopyIdentity(const hepem_tracks& in, hepem_tracks& out, int i) {
  if (out.ekin == in.ekin) return;
  out.ptype[i] = in.ptype[i];
  out.id[i] = in.id[i];
  if (out.rng_stream) out.rng_stream[i] = StreamHi(in, i);
  out.imc[i] = in.imc[i];
  out.safety[i] = in.safety[i];
  out.on_boundary[i] = in.on_boundary[i];
}

HD void StoreRng(hepem_tracks& out, int i, const Rng& rng) {
  out.rng_ctr[i] = rng.ctr;
  out.rng_gauss[i] = rng.gauss;
  out.rng_has_gauss[i] = rng.has_gauss ? 1 : 0;
}

// Loads the start-of-step state of slot i common to both particle kinds.
template <typename R> HD void LoadTrack(const hepem_tracks& in, int n, int i, Trk<R>& t) {
  t.setE(in.ekin[i]);
#pragma unroll
  for (int c = 0; c < 3; ++c) t.dir[c] = (R)in.dir[c * n + i];
  t.imc = in.imc[i];
  t.safety = (R)in.safety[i];
  t.onBoundary = in.on_boundary[i] != 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    t.nial[k] = (R)in.num_ia_left[k * n + i];
    t.mfp[k] = (R)-1.0;
  }
  t.winner = -1;
  t.edep = 0.0;
  t.gstep = (R)0.0;
}

// Start-of-step state of an e-/e+ slot (G4HepEmElectronTrack::ReSet with the
// persistent MSC state and the RNG state of the slot).
template <typename R>
HD void InitElectron(const hepem_tracks& in, int n, int i, uint64_t seed, ElState<R>& el, Rng& rng) {
  Trk<R>& t = el.t;
  t.isEl = in.ptype[i] == 0;
  LoadTrack(in, n, i, t);
  Msc<R>& m = el.msc;
  m.lambtr1 = m.trueStep = m.zpath = (R)0.;
  m.displ[0] = m.displ[1] = m.displ[2] = (R)0.;
  m.dir[0] = m.dir[1] = (R)0.;
  m.dir[2] = (R)1.;
  m.initRange = (R)in.msc_init_range[i];
  m.dynRangeFactor = (R)in.msc_dyn_range_factor[i];
  m.tlimitMin = (R)in.msc_tlimit_min[i];
  m.par1 = (R)-1.;
  m.par2 = m.par3 = (R)0.;
  m.noScat = m.isDisplace = m.active = false;
  m.firstStep = in.msc_first_step[i] != 0;
  el.range = el.pstep = (R)0.;
  rng.init(seed, (uint32_t)in.id[i], StreamHi(in, i), in.rng_ctr[i], in.rng_has_gauss[i] != 0, in.rng_gauss[i]);
}

// G4HepEmElectronManager::Perform (G4HepEmElectronManager.icc:461-483).
template <typename R> HD void ElectronPerform(const DevTables<R>& T, ElState<R>& el, SecList<R>& sl, Rng& rng) {
  Trk<R>& t = el.t;
  const int cont = ElectronPerformContinuous(T, el, rng);
  if (cont == 1) {
    if (!t.isEl) AnnihilateAtRest(sl, rng);
  } else if (cont == 0) {
    // PerformDiscrete (:425-459)
    switch (ElectronSelectDiscrete(T, t, rng)) {
      case 0: IoniPerform(T, t, sl, rng); break;
      case 1: BremPerform(T, t, sl, rng, t.ekin < T.bremModelLim); break;
      case 2: AnnihilateInFlight(t, sl, rng); break;
      default: break;
    }
  }
}

// Persistent state (interaction lengths, MSC, RNG) of an e-/e+ slot.
template <typename R> HD void StoreElectronState(hepem_tracks& out, int n, int i, const ElState<R>& el, const Rng& rng) {
  const Trk<R>& t = el.t;
  const Msc<R>& m = el.msc;
#pragma unroll
  for (int k = 0; k < 4; ++k) out.num_ia_left[k * n + i] = (double)t.nial[k];
  out.msc_init_range[i] = (double)m.initRange;
  out.msc_dyn_range_factor[i] = (double)m.dynRangeFactor;
  out.msc_tlimit_min[i] = (double)m.tlimitMin;
  out.msc_first_step[i] = m.firstStep ? 1 : 0;
  StoreRng(out, i, rng);
}

// Post-step kinematics and step results of an e-/e+ slot.
template <typename R>
HD void StoreElectronResult(hepem_tracks& out, const hepem_step_out& res, int n, int i, const ElState<R>& el,
                            const SecList<R>& sl) {
  const Trk<R>& t = el.t;
  out.ekin[i] = t.ekin;
#pragma unroll
  for (int c = 0; c < 3; ++c) out.dir[c * n + i] = (double)t.dir[c];
  res.true_step[i] = (double)el.pstep;
  res.geom_step[i] = (double)t.gstep;
  res.edep[i] = t.edep;
#pragma unroll
  for (int c = 0; c < 3; ++c) res.displacement[c * n + i] = (double)el.msc.displ[c];
  res.winner[i] = t.winner;
  res.stopped[i] = t.ekin <= 0.0 ? 1 : 0;
  WriteSecondaries(res, n, i, sl);
}

// HowFar -> Perform hand-over of an e-/e+ slot.
template <typename R> HD void StoreElectronHowFar(const hepem_howfar_state& hs, int n, int i, const ElState<R>& el) {
  const Trk<R>& t = el.t;
  const Msc<R>& m = el.msc;
  hs.geom_step[i] = (double)t.gstep;
  hs.range[i] = (double)el.range;
#pragma unroll
  for (int k = 0; k < 4; ++k) hs.mfp[k * n + i] = (double)t.mfp[k];
  hs.winner[i] = t.winner;
  hs.msc_lambtr1[i] = (double)m.lambtr1;
  hs.msc_true_step[i] = (double)m.trueStep;
  hs.msc_zpath[i] = (double)m.zpath;
  hs.msc_par1[i] = (double)m.par1;
  hs.msc_par2[i] = (double)m.par2;
  hs.msc_par3[i] = (double)m.par3;
  hs.msc_active[i] = m.active ? 1 : 0;
  hs.msc_displace[i] = m.isDisplace ? 1 : 0;
}
template <typename R> HD void LoadElectronHowFar(const hepem_howfar_state& hs, int n, int i, ElState<R>& el) {
  Trk<R>& t = el.t;
  Msc<R>& m = el.msc;
  t.gstep = (R)hs.geom_step[i];
  el.range = (R)hs.range[i];
#pragma unroll
  for (int k = 0; k < 4; ++k) t.mfp[k] = (R)hs.mfp[k * n + i];
  t.winner = hs.winner[i];
  m.lambtr1 = (R)hs.msc_lambtr1[i];
  m.trueStep = (R)hs.msc_true_step[i];
  m.zpath = (R)hs.msc_zpath[i];
  m.par1 = (R)hs.msc_par1[i];
  m.par2 = (R)hs.msc_par2[i];
  m.par3 = (R)hs.msc_par3[i];
  m.active = hs.msc_active[i] != 0;
  m.isDisplace = hs.msc_displace[i] != 0;
}

// G4HepEmElectronManager::HowFar + Perform (G4HepEmElectronManager.icc:35-45,
// 461-483) of one e-/e+ slot, no geometry in between.
template <typename R>
HD void StepElectron(const DevTables<R>& T, uint64_t seed, const hepem_tracks& in, hepem_tracks& out, const hepem_step_out& res,
                     int n, int i) {
  ElState<R> el;
  Rng rng;
  InitElectron(in, n, i, seed, el, rng);
  SecList<R> sl;
  sl.ns = 0;
  ElectronHowFar(T, el, rng);
  ElectronPerform(T, el, sl, rng);
  StoreElectronState(out, n, i, el, rng);
  StoreElectronResult(out, res, n, i, el, sl);
}

// G4HepEmGammaManager::HowFar (G4HepEmGammaManager.icc:27-48): returns the PE
// macroscopic cross section stored in the gamma track.
template <typename R> HD R GammaHowFar(const DevTables<R>& T, Trk<R>& t, Rng& rng) {
  R pe = (R)0.0;
  if (t.nial[0] <= (R)0.0) t.nial[0] = -vlog((R)rng.flat());
  const R tot = GetTotalMacXSec(T, t, pe);
  const R mfp = (tot > (R)0.) ? (R)1. / tot : (R)kALargeValue;
  t.mfp[0] = mfp;
  t.gstep = mfp * t.nial[0];
  return pe;
}

// What follows the gamma's move (G4HepEmTrackingManager.cc:1092-1108): on a
// boundary only the interaction-length update, otherwise SelectInteraction and
// Perform (G4HepEmGammaManager.icc:54-94,173-219).
template <typename R> HD void GammaPerform(const DevTables<R>& T, Trk<R>& t, R pe, SecList<R>& sl, Rng& rng) {
  if (t.onBoundary) {
    t.nial[0] -= t.gstep / t.mfp[0];
    return;
  }
  SampleInteraction(T, t, pe, (R)rng.flat());
  if (t.winner == 3) return;
  t.nial[0] -= t.gstep / t.mfp[0];
  ResetNumIALeft(t.nial, t.winner);
  switch (t.winner) {
    case 0: ConvPerform(T, t, sl, rng); break;
    case 1: ComptonPerform(t, sl, rng); break;
    case 2: PEPerform(T, t, pe, sl, rng); break;
    default: break;
  }
  const double finalEkin = t.ekin;
  if (finalEkin > 0.0 && finalEkin <= T.gamTrackCut) {
    t.setE(0.0);
    t.edep += finalEkin;
  }
}

template <typename R>
HD void StoreGamma(hepem_tracks& out, const hepem_step_out& res, int n, int i, const Trk<R>& t, const SecList<R>& sl,
                   const Rng& rng) {
  out.ekin[i] = t.ekin;
#pragma unroll
  for (int c = 0; c < 3; ++c) out.dir[c * n + i] = (double)t.dir[c];
#pragma unroll
  for (int k = 0; k < 4; ++k) out.num_ia_left[k * n + i] = (double)t.nial[k];
  StoreRng(out, i, rng);
  res.true_step[i] = (double)t.gstep;
  res.geom_step[i] = (double)t.gstep;
  res.edep[i] = t.edep;
#pragma unroll
  for (int c = 0; c < 3; ++c) res.displacement[c * n + i] = 0.0;
  res.winner[i] = t.winner;
  res.stopped[i] = t.ekin <= 0.0 ? 1 : 0;
  WriteSecondaries(res, n, i, sl);
}

// G4HepEmGammaManager::HowFar + SelectInteraction + Perform of one gamma slot,
// no geometry in between.
template <typename R>
HD void StepGamma(const DevTables<R>& T, uint64_t seed, const hepem_tracks& in, hepem_tracks& out, const hepem_step_out& res,
                  int n, int i) {
  Trk<R> t;
  t.isEl = false;
  LoadTrack(in, n, i, t);
  Rng rng;
  rng.init(seed, (uint32_t)in.id[i], StreamHi(in, i), in.rng_ctr[i], in.rng_has_gauss[i] != 0, in.rng_gauss[i]);
  SecList<R> sl;
  sl.ns = 0;
  const R pe = GammaHowFar(T, t, rng);
  GammaPerform(T, t, pe, sl, rng);
  StoreGamma(out, res, n, i, t, sl, rng);
}

}  // namespace hepem
