// dev_tracking.hh -- one e-/e+ step as G4HepEmTrackingManager::TrackElectron
// takes it (G4HepEm/G4HepEm/src/G4HepEmTrackingManager.cc:408-635), in phases
// around the caller's geometry:
//   LoopBegin   the interaction lengths and HowFarToDiscreteInteraction, once
//               per step (:427-443);
//   LoopMsc     HowFarToMSC of a sub-step (:446-452); the caller then moves the
//               track at most the proposed geometrical length and reports
//               whether a boundary cut the move (:456-486);
//   LoopAlong   UpdatePStepLength, UpdateNumIALeft, ApplyMeanEnergyLoss and
//               SampleMSC of the sub-step (:490-521); the caller applies the
//               MSC displacement (:523-566); while MSC limited the sub-step,
//               the region sets fIsMultipleStepsInMSCTrans and no boundary was
//               reached, the next sub-step is what is left of the discrete step
//               (:571-594);
//   LoopFinish  the loss fluctuation on the summed mean loss from the step's
//               pre-step energy (:597-606), then at-rest annihilation or
//               PerformDiscrete (:611-650; winner 3 resets its interaction
//               length and draws the delta test, as with a lepto-nuclear process
//               attached).
// ElectronSlabStep runs them with the caller's geometry a slab across x.
#pragma once

#include <cfloat>

#include "dev_physics.hh"

namespace hepem {

constexpr int kMaxSubsteps = 65536;  // the reference's sub-step loop is unbounded

// The locals of TrackElectron that live across the sub-steps of one step.
template <typename R> struct StepLoop {
  int iDProc;        // the discrete process HowFarToDiscreteInteraction chose
  R stepLeft;        // stepLimitLeft: what is left of the discrete step
  R truePath;        // totalTruePathLength
  double meanLoss;   // totalEloss
  double preEkin;    // the step's pre-step energy and its logarithm
  R preLEkin;
  bool cont;         // continueStepping
  bool stopped;
  bool left;         // the track left the world (set by the caller's geometry)
  int nsub;
};

template <typename R> HD void LoopBegin(const DevTables<R>& T, ElState<R>& el, StepLoop<R>& L, Rng& rng) {
  Trk<R>& t = el.t;
  L.preEkin = t.ekin;
  L.preLEkin = t.logE();
  L.cont = T.regMultiStep[T.mcReg[t.imc]] != 0;
#pragma unroll
  for (int ip = 0; ip < 4; ++ip)
    if (t.nial[ip] <= (R)0.) t.nial[ip] = -vlog((R)rng.flat());
  HowFarToDiscreteInteraction(T, el);
  L.iDProc = t.winner;
  L.stepLeft = el.pstep;
  L.truePath = (R)0.;
  L.meanLoss = 0.0;
  L.stopped = false;
  L.left = false;
  L.nsub = 0;
  el.preEkin = t.ekin;  // SavePreStepEKin
  el.preLEkin = t.logE();
}

// HowFarToMSC of the next sub-step; t.gstep is then the proposed geometrical length.
template <typename R> HD void LoopMsc(const DevTables<R>& T, ElState<R>& el, StepLoop<R>& L, Rng& rng) {
  ++L.nsub;
  HowFarToMSC(T, el, rng);
  if (el.t.winner != -2) L.cont = false;
}

// The sub-step after the caller's move: t.gstep = the length moved, t.onBoundary
// = the move ended on a boundary (which ends the loop). Returns whether the MSC
// scattering was sampled; el.msc.displ is then the displacement to apply off a
// boundary (zero if none was sampled).
template <typename R> HD bool LoopAlong(const DevTables<R>& T, ElState<R>& el, StepLoop<R>& L, Rng& rng) {
  Trk<R>& t = el.t;
  if (t.onBoundary) L.cont = false;
  if (!(t.gstep > (R)0)) return false;
  UpdatePStepLength(el);
  const R pStep = el.pstep;
  L.truePath += pStep;
  bool sampled = false;
  if (pStep > (R)0.0) {
    UpdateNumIALeft(el);
    L.stopped = ApplyMeanEnergyLoss(T, el);
    L.meanLoss += t.edep;
    if (L.stopped) {
      L.cont = false;
    } else {
      SampleMSC(T, el, rng);
      sampled = true;
    }
  }
  if (L.cont) {
    t.edep = 0.0;
    t.winner = L.iDProc;
    el.preEkin = t.ekin;
    el.preLEkin = t.logE();
    L.stepLeft -= pStep;
    el.pstep = L.stepLeft;
    t.gstep = L.stepLeft;
    el.range -= pStep;
  }
  return sampled;
}

// LoopFinish up to the final-state sampling of the discrete interaction:
// returns the process whose final state is to be sampled (0 ionisation,
// 1 brem, 2 annihilation in flight; ElectronFinalState) or -1.
template <typename R>
HD int LoopFinishSelect(const DevTables<R>& T, ElState<R>& el, StepLoop<R>& L, SecList<R>& sl, Rng& rng) {
  Trk<R>& t = el.t;
  t.edep = L.meanLoss;
  if (!L.stopped) {
    el.preEkin = L.preEkin;
    el.preLEkin = L.preLEkin;
    L.stopped = SampleLossFluctuations(T, el, rng);
  }
  if (L.stopped) {
    if (!t.isEl) AnnihilateAtRest(sl, rng);
    return -1;
  }
  if (L.left) return -1;
  const int proc = ElectronSelectDiscrete(T, t, rng);
  return proc >= 0 && proc < 3 ? proc : -1;
}

// The final state of the discrete interaction `proc` (PerformDiscrete,
// G4HepEmElectronManager.icc:443-458).
template <typename R> HD void ElectronFinalState(const DevTables<R>& T, Trk<R>& t, int proc, SecList<R>& sl, Rng& rng) {
  switch (proc) {
    case 0: IoniPerform(T, t, sl, rng); break;
    case 1: BremPerform(T, t, sl, rng, t.ekin < T.bremModelLim); break;
    case 2: AnnihilateInFlight(t, sl, rng); break;
    default: break;
  }
}

template <typename R> HD void LoopFinish(const DevTables<R>& T, ElState<R>& el, StepLoop<R>& L, SecList<R>& sl, Rng& rng) {
  ElectronFinalState(T, el.t, LoopFinishSelect(T, el, L, sl, rng), sl, rng);
}

// The MSC displacement of a sub-step that did not end on a boundary
// (G4HepEmTrackingManager.cc:527-566), with the safety of a point x in a slab
// lo <= x <= hi across x: its distance to the nearer plane.
template <typename R> HD double DisplaceInSlab(const Msc<R>& m, double x, double lo, double hi) {
  const double d0 = (double)m.displ[0], d1 = (double)m.displ[1], d2 = (double)m.displ[2];
  const double dl2 = d0 * d0 + d1 * d1 + d2 * d2;
  constexpr double kMin = 5.0e-8;
  if (dl2 > kMin * kMin) {
    const double dispR = sqrt(dl2);
    const double ps = 0.99 * fmax(fmin(x - lo, hi - x), 0.0);
    if (ps > 0.0 && dispR <= ps) {
      x += d0;
    } else if (dispR < ps) {
      x += d0;
    } else if (ps > kMin) {
      x += d0 * (ps / dispR);
    }
  }
  return x;
}

// One TrackElectron step of a track at x in the slab lo <= x <= hi across x;
// crossing lo (exitLo) or hi (exitHi) leaves the world. The pre-step safety is
// the distance to the nearer plane, 0 on a boundary. On return x is the
// post-step position, side +1/-1 if the step ended on hi/lo (else 0) and
// L.left whether the track left the world.
struct SlabPos {
  double x, lo, hi;
  bool exitLo, exitHi;
  int side;
};

// The slab step in pieces, for callers that interleave the sub-steps of
// several tracks: SlabBegin, then SlabSubstep until it returns true, then
// LoopFinishSelect (or LoopFinish).
template <typename R> HD void SlabBegin(const DevTables<R>& T, ElState<R>& el, StepLoop<R>& L, Rng& rng, SlabPos& g) {
  Trk<R>& t = el.t;
  t.safety = t.onBoundary ? (R)0.0 : (R)fmax(fmin(g.x - g.lo, g.hi - g.x), 0.0);
  LoopBegin(T, el, L, rng);
  g.side = 0;
}

// One sub-step: HowFarToMSC, the move to the next plane at most, the along-step
// part and the MSC displacement. Returns true when the step's sub-steps are done.
template <typename R> HD bool SlabSubstep(const DevTables<R>& T, ElState<R>& el, StepLoop<R>& L, Rng& rng, SlabPos& g) {
  Trk<R>& t = el.t;
  double x = g.x;
  LoopMsc(T, el, L, rng);
  const double physicalStep = (double)t.gstep;
  const double ux = (double)t.dir[0];
  double dist = DBL_MAX;
  int sd = 0;
  if (ux > 0.0) {
    dist = (g.hi - x) / ux;
    sd = 1;
  } else if (ux < 0.0) {
    dist = (g.lo - x) / ux;
    sd = -1;
  }
  const bool limited = dist < physicalStep;
  const double finalStep = limited ? dist : physicalStep;
  if (limited) {
    x = sd > 0 ? g.hi : g.lo;
    g.side = sd;
    if ((sd > 0 && g.exitHi) || (sd < 0 && g.exitLo)) {
      L.left = true;
      L.cont = false;
      g.x = x;
      return true;
    }
  } else {
    x = fmin(fmax(x + finalStep * ux, g.lo), g.hi);
  }
  t.gstep = (R)finalStep;
  t.onBoundary = limited;
  if (LoopAlong(T, el, L, rng) && !limited) x = DisplaceInSlab(el.msc, x, g.lo, g.hi);
  g.x = x;
  return !L.cont || L.nsub >= kMaxSubsteps;
}

// ElectronSlabStep up to the final-state sampling: returns the process whose
// final state is to be sampled (see LoopFinishSelect) or -1.
template <typename R>
HD int ElectronSlabStepSelect(const DevTables<R>& T, ElState<R>& el, StepLoop<R>& L, SecList<R>& sl, Rng& rng, SlabPos& g) {
  SlabBegin(T, el, L, rng, g);
  while (!SlabSubstep(T, el, L, rng, g)) {
  }
  return LoopFinishSelect(T, el, L, sl, rng);
}

template <typename R>
HD void ElectronSlabStep(const DevTables<R>& T, ElState<R>& el, StepLoop<R>& L, SecList<R>& sl, Rng& rng, SlabPos& g) {
  ElectronFinalState(T, el.t, ElectronSlabStepSelect(T, el, L, sl, rng, g), sl, rng);
}

}  // namespace hepem
