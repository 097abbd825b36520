// hepem_kernels.hh -- kernels of the stepping library (one translation unit,
// included by hepem_capi.hip).
//
// Stepping is split in two launches:
//   k_classify  compaction of track slots by particle type into one queue:
//               e-/e+ slots fill it from the front, gamma slots from the back
//               (wave ballots and an LDS scan per tile of kClassifyTile slots,
//               one atomic per tile and kind);
//   k_step      one thread per queue entry; the e-/e+ and gamma paths are
//               therefore taken by whole waves (only the one wave that straddles
//               the boundary diverges), and the long e-/e+ blocks are dealt out
//               before the short gamma blocks.
// Every track writes only its own slot, so results do not depend on the queue
// order and are bit-identical to an unsorted pass.
#pragma once

#include "device/dev_physics.hh"
#include "device/dev_soa.hh"
#include "device/dev_tracking.hh"
#include "g4hepem_amd.h"

namespace hepem {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

// counters[0]: e-/e+ tracks, counters[1]: gammas, counters[2]: rejected tracks,
// counters[3]: 1 when the queue holds the tile-local partition (counters[4..63]
// reserved)
constexpr int kCounterWords = 64;
constexpr int kCounterLayout = 3;
__global__ void k_fill_u32(uint32_t* p, int n, uint32_t v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Each block compacts a tile of kClassifyTile consecutive slots in two
// passes over the tile: count per wave (ballot + popcount), one atomic per
// block and kind for the tile's base, then scatter. The tile keeps atomic
// traffic on the two queue counters to a few hundred operations per launch.
constexpr int kClassifyBlock = 1024;  // threads of the classification kernels
constexpr int kClassifyWaves = kClassifyBlock / 64;
constexpr int kClassifyItems = 4;
constexpr int kClassifyTile = kClassifyBlock * kClassifyItems;

// Compacts tile `tile` of a batch of n slots into its queue (e-/e+ from the
// front, gammas from the back); counters[0..1] count e-/e+ and gammas.
__device__ void ClassifyTile(int n, const int8_t* __restrict__ ptype, uint32_t* __restrict__ queue,
                             uint32_t* __restrict__ counters, int tile) {
  __shared__ uint32_t sE[kClassifyWaves], sG[kClassifyWaves], sBase[2];
  const int lane = __lane_id();
  const int wave = threadIdx.x >> 6;
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int tile0 = tile * kClassifyTile;
  uint32_t cE = 0, cG = 0;
  for (int it = 0; it < kClassifyItems; ++it) {
    const int i = tile0 + it * kClassifyBlock + threadIdx.x;
    const bool valid = i < n;
    const int8_t pt = valid ? ptype[i] : (int8_t)-1;
    const bool isE = valid && (pt == 0 || pt == 1);
    cE += (uint32_t)__popcll(__ballot(isE));
    cG += (uint32_t)__popcll(__ballot(valid && !isE));
  }
  if (lane == 0) {
    sE[wave] = cE;
    sG[wave] = cG;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tE = 0, tG = 0;
    for (int w = 0; w < kClassifyWaves; ++w) {
      const uint32_t e = sE[w], g = sG[w];
      sE[w] = tE;
      sG[w] = tG;
      tE += e;
      tG += g;
    }
    sBase[0] = tE ? atomicAdd(&counters[0], tE) : 0u;
    sBase[1] = tG ? atomicAdd(&counters[1], tG) : 0u;
  }
  __syncthreads();
  uint32_t pE = sBase[0] + sE[wave], pG = sBase[1] + sG[wave];
  for (int it = 0; it < kClassifyItems; ++it) {
    const int i = tile0 + it * kClassifyBlock + threadIdx.x;
    const bool valid = i < n;
    const int8_t pt = valid ? ptype[i] : (int8_t)-1;
    const bool isE = valid && (pt == 0 || pt == 1);
    const bool isG = valid && !isE;
    const uint64_t mE = __ballot(isE), mG = __ballot(isG);
    if (isE) queue[pE + __popcll(mE & below)] = (uint32_t)i;
    if (isG) queue[(uint32_t)n - 1u - (pG + __popcll(mG & below))] = (uint32_t)i;
    pE += (uint32_t)__popcll(mE);
    pG += (uint32_t)__popcll(mG);
  }
}

__global__ void __launch_bounds__(kClassifyBlock) k_classify(int n, const int8_t* __restrict__ ptype, uint32_t* __restrict__ queue,
                                                     uint32_t* __restrict__ counters) {
  ClassifyTile(n, ptype, queue, counters, blockIdx.x);
}

// kinds of slots a kernel steps: a queue-ordered mixed batch, or a batch of
// one kind (other slots rejected)
enum SlotKind : int { kMixed = 0, kOnlyElectrons = 1, kOnlyGammas = 2 };

// Resolves thread j of a stepping kernel to (slot, is-e-/e+); false if the
// thread has no slot or the slot does not hold a track of the batch's kind.
template <int KIND>
HD bool ResolveSlot(const hepem_tracks& t, int j, const uint32_t* queue, const uint32_t* counters, int& i, bool& isE) {
  if (KIND == kMixed) {
    i = (int)queue[j];
    isE = (uint32_t)j < counters[0];
    return true;
  }
  i = j;
  const int8_t pt = t.ptype[i];
  isE = pt == 0 || pt == 1;
  return KIND == kOnlyElectrons ? isE : pt == 2;
}

// ---- TrackElectron's sub-step loop driven by the caller (device/dev_tracking.hh) ----------------
// counters[4]: tracks that take another sub-step after the last begin/along call

// Adds the number of lanes of the wave with `pred` to *c (one atomic per wave);
// every lane of the wave must call it.
__device__ inline void CountWave(uint32_t* c, bool pred) {
  const uint64_t m = __ballot(pred);
  if (m && __lane_id() == __ffsll((unsigned long long)m) - 1) atomicAdd(c, (uint32_t)__popcll(m));
}

template <typename R> HD void LoadLoop(const hepem_substep_state& ss, int i, StepLoop<R>& L) {
  L.iDProc = ss.discrete_winner[i];
  L.stepLeft = (R)ss.step_left[i];
  L.truePath = (R)ss.true_path[i];
  L.meanLoss = ss.mean_eloss[i];
  L.preEkin = ss.pre_ekin[i];
  L.preLEkin = L.preEkin > 0. ? vlog((R)L.preEkin) : (R)-30;
  L.cont = ss.more[i] != 0;
  L.stopped = ss.stopped[i] != 0;
  L.left = ss.left_world[i] != 0;
  L.nsub = ss.nsub[i];
}
template <typename R> HD void StoreLoop(const hepem_substep_state& ss, int i, const StepLoop<R>& L) {
  ss.discrete_winner[i] = L.iDProc;
  ss.step_left[i] = (double)L.stepLeft;
  ss.true_path[i] = (double)L.truePath;
  ss.mean_eloss[i] = L.meanLoss;
  ss.pre_ekin[i] = L.preEkin;
  ss.more[i] = L.cont ? 1 : 0;
  ss.stopped[i] = L.stopped ? 1 : 0;
  ss.nsub[i] = L.nsub;
}

template <typename R>
__global__ void __launch_bounds__(kBlock) k_el_begin(DevTables<R> T, uint64_t seed, hepem_tracks t, hepem_howfar_state hs,
                                                     hepem_substep_state ss, uint32_t* __restrict__ counters) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = t.n;
  bool active = false;
  if (i < n) {
    const int8_t pt = t.ptype[i];
    if ((pt != 0 && pt != 1) || !TrackValid(T, t, i)) {
      atomicAdd(&counters[2], 1u);
      hs.winner[i] = -99;
      hs.geom_step[i] = 0.0;
      ss.active[i] = 0;
      ss.more[i] = 0;
    } else {
      ElState<R> el;
      Rng rng;
      InitElectron(t, n, i, seed, el, rng);
      StepLoop<R> L;
      LoopBegin(T, el, L, rng);
      StoreElectronState(t, n, i, el, rng);
      StoreElectronHowFar(hs, n, i, el);
      StoreLoop(ss, i, L);
      ss.geom_path[i] = 0.0;
#pragma unroll
      for (int c = 0; c < 3; ++c) ss.displacement[c * n + i] = 0.0;
      ss.active[i] = 1;
      ss.left_world[i] = 0;
      active = true;
    }
  }
  CountWave(&counters[4], active);
}

template <typename R>
__global__ void __launch_bounds__(kBlock) k_el_substep_howfar(DevTables<R> T, uint64_t seed, hepem_tracks t,
                                                              hepem_howfar_state hs, hepem_substep_state ss) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = t.n;
  if (i >= n || !ss.active[i]) return;
  ElState<R> el;
  Rng rng;
  InitElectron(t, n, i, seed, el, rng);
  LoadElectronHowFar(hs, n, i, el);
  el.pstep = (R)ss.step_left[i];
  StepLoop<R> L;
  LoadLoop(ss, i, L);
  LoopMsc(T, el, L, rng);
  StoreElectronState(t, n, i, el, rng);
  StoreElectronHowFar(hs, n, i, el);
  ss.more[i] = L.cont ? 1 : 0;
  ss.nsub[i] = L.nsub;
}

template <typename R>
__global__ void __launch_bounds__(kBlock) k_el_substep_along(DevTables<R> T, uint64_t seed, hepem_tracks t,
                                                             hepem_howfar_state hs, hepem_substep_state ss,
                                                             uint32_t* __restrict__ counters) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = t.n;
  bool again = false;
  if (i < n && ss.active[i]) {
    ss.geom_path[i] += hs.geom_step[i];
    if (ss.left_world[i]) {
      ss.active[i] = 0;
      ss.more[i] = 0;
    } else {
      ElState<R> el;
      Rng rng;
      InitElectron(t, n, i, seed, el, rng);
      LoadElectronHowFar(hs, n, i, el);
      el.preEkin = el.t.ekin;
      el.preLEkin = el.t.logE();
      StepLoop<R> L;
      LoadLoop(ss, i, L);
      const bool sampled = LoopAlong(T, el, L, rng);
      StoreElectronState(t, n, i, el, rng);
      t.ekin[i] = el.t.ekin;
#pragma unroll
      for (int c = 0; c < 3; ++c) t.dir[c * n + i] = (double)el.t.dir[c];
      StoreElectronHowFar(hs, n, i, el);
      StoreLoop(ss, i, L);
#pragma unroll
      for (int c = 0; c < 3; ++c) ss.displacement[c * n + i] = sampled ? (double)el.msc.displ[c] : 0.0;
      again = L.cont;
      ss.active[i] = again ? 1 : 0;
    }
  }
  CountWave(&counters[4], again);
}

template <typename R>
__global__ void __launch_bounds__(kBlock) k_el_end(DevTables<R> T, uint64_t seed, hepem_tracks t, hepem_howfar_state hs,
                                                   hepem_substep_state ss, hepem_step_out res) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = t.n;
  if (i >= n) return;
  const int8_t pt = t.ptype[i];
  if ((pt != 0 && pt != 1) || hs.winner[i] == -99) {
    WriteRejected<R>(t, t, res, n, i);
    return;
  }
  ElState<R> el;
  Rng rng;
  InitElectron(t, n, i, seed, el, rng);
  LoadElectronHowFar(hs, n, i, el);
  StepLoop<R> L;
  LoadLoop(ss, i, L);
  SecList<R> sl;
  sl.ns = 0;
  LoopFinish(T, el, L, sl, rng);
  StoreElectronState(t, n, i, el, rng);
  t.ekin[i] = el.t.ekin;
#pragma unroll
  for (int c = 0; c < 3; ++c) t.dir[c * n + i] = (double)el.t.dir[c];
  ss.stopped[i] = L.stopped ? 1 : 0;
  res.true_step[i] = (double)L.truePath;
  res.geom_step[i] = ss.geom_path[i];
  res.edep[i] = el.t.edep;
#pragma unroll
  for (int c = 0; c < 3; ++c) res.displacement[c * n + i] = 0.0;
  res.winner[i] = el.t.winner;
  res.stopped[i] = el.t.ekin <= 0.0 ? 1 : 0;
  WriteSecondaries(res, n, i, sl);
}

// ---- two-stage step: particle sort, then process sort ----------------------------------------
// k_step_select_local does HowFar, the continuous step and the choice of the
// discrete interaction (one thread per particle-queue entry), records the chosen process
// per slot and, for a pending interaction, the 64-byte record of what the
// final-state sampling needs (at the slot: one line per track instead of ten
// scattered fields); k_classify_tiles orders each tile's pending slots by
// process; k_interact_tiles samples the final states, so each wave runs one
// sampler. (The batched step does the same over the tiles of all its batches,
// k_classify_tiles_batched and k_interact_tiles_batched.)

// processes of the second stage (0: none pending), in the low 4 bits of the
// per-slot code; bit 4: the track is an e-, bit 5: a Gaussian deviate is cached
enum ProcCode : uint8_t {
  kPNone = 0,
  kPIoniEl,    // Moller
  kPIoniPos,   // Bhabha
  kPBremSB,    // Seltzer-Berger brem (E < model limit)
  kPBremRB,    // relativistic brem
  kPAnnih,     // e+ annihilation in flight
  kPConv,      // gamma conversion
  kPCompton,   // Klein-Nishina
  kPPE,        // photoelectric
};
constexpr int kNumProc = 8;          // codes 1..8
constexpr uint8_t kCodeProc = 0x0F, kCodeElectron = 0x10, kCodeGauss = 0x20;

// state handed from the selection to the final-state sampling of one slot
struct alignas(16) Staged {
  double ekin, dir[3], gauss, pe;
  int32_t imc, id;
  uint32_t stream, ctr;
};
static_assert(sizeof(Staged) == 64, "one 64-byte record per pending slot");

HD uint8_t StageTrack(Staged* st, int i, uint8_t proc, bool isEl, int imc, int32_t id, uint32_t stream, double ekin,
                      const double* dir, const Rng& rng, double pe) {
  Staged r;
  r.ekin = ekin;
  r.dir[0] = dir[0];
  r.dir[1] = dir[1];
  r.dir[2] = dir[2];
  r.gauss = rng.gauss;
  r.pe = pe;
  r.imc = imc;
  r.id = id;
  r.stream = stream;
  r.ctr = rng.ctr;
  st[i] = r;
  return (uint8_t)(proc | (isEl ? kCodeElectron : 0) | (rng.has_gauss ? kCodeGauss : 0));
}

// Stage 1 of an e-/e+ slot after its HowFar: the continuous step and
// PerformDiscrete's choice (G4HepEmElectronManager.icc:425-441,461-483).
// Returns the slot code; a complete step (kPNone) also gets its post-step
// kinematics, secondaries and stopped flag here, a pending one its staged
// record.
template <typename R>
HD uint8_t ElectronSelectAfterHowFar(const DevTables<R>& T, ElState<R>& el, Rng& rng, const hepem_tracks& in,
                                     hepem_tracks& out, const hepem_step_out& res, Staged* st, int n, int i) {
  Trk<R>& t = el.t;
  SecList<R> sl;
  sl.ns = 0;
  uint8_t proc = kPNone;
  const int cont = ElectronPerformContinuous(T, el, rng);
  if (cont == 1) {
    if (!t.isEl) AnnihilateAtRest(sl, rng);
  } else if (cont == 0) {
    const int ip = ElectronSelectDiscrete(T, t, rng);
    proc = ip == 0   ? (t.isEl ? kPIoniEl : kPIoniPos)
           : ip == 1 ? (t.ekin < T.bremModelLim ? kPBremSB : kPBremRB)
           : ip == 2 ? kPAnnih
                     : kPNone;  // -1, or e-/e+ nuclear (no final state)
  }
  const Msc<R>& m = el.msc;
#pragma unroll
  for (int k = 0; k < 4; ++k) out.num_ia_left[k * n + i] = (double)t.nial[k];
  out.msc_init_range[i] = (double)m.initRange;
  out.msc_dyn_range_factor[i] = (double)m.dynRangeFactor;
  out.msc_tlimit_min[i] = (double)m.tlimitMin;
  out.msc_first_step[i] = m.firstStep ? 1 : 0;
  res.true_step[i] = (double)el.pstep;
  res.geom_step[i] = (double)t.gstep;
  res.edep[i] = t.edep;
#pragma unroll
  for (int c = 0; c < 3; ++c) res.displacement[c * n + i] = (double)m.displ[c];
  res.winner[i] = t.winner;
  if (proc != kPNone) {
    const double d[3] = {(double)t.dir[0], (double)t.dir[1], (double)t.dir[2]};
    return StageTrack(st, i, proc, t.isEl, t.imc, in.id[i], StreamHi(in, i), t.ekin, d, rng, 0.0);
  }
  out.ekin[i] = t.ekin;
#pragma unroll
  for (int c = 0; c < 3; ++c) out.dir[c * n + i] = (double)t.dir[c];
  StoreRng(out, i, rng);
  res.stopped[i] = t.ekin <= 0.0 ? 1 : 0;
  WriteSecondaries(res, n, i, sl);
  return kPNone;
}

// Stage 1 of an e-/e+ slot of the fused step: as StepElectron up to
// PerformDiscrete's choice.
template <typename R>
HD uint8_t StepElectronSelect(const DevTables<R>& T, uint64_t seed, const hepem_tracks& in, hepem_tracks& out,
                              const hepem_step_out& res, Staged* st, int n, int i) {
  ElState<R> el;
  Rng rng;
  InitElectron(in, n, i, seed, el, rng);
  ElectronHowFar(T, el, rng);
  return ElectronSelectAfterHowFar(T, el, rng, in, out, res, st, n, i);
}

// Stage 1 of an e-/e+ slot of the split Perform: the HowFar state comes from
// the hand-over (with the caller's geometrical step).
template <typename R>
HD uint8_t PerformElectronSelect(const DevTables<R>& T, uint64_t seed, hepem_tracks& t, const hepem_howfar_state& hs,
                                 const hepem_step_out& res, Staged* st, int n, int i) {
  ElState<R> el;
  Rng rng;
  InitElectron(t, n, i, seed, el, rng);
  LoadElectronHowFar(hs, n, i, el);
  return ElectronSelectAfterHowFar(T, el, rng, t, t, res, st, n, i);
}

// Stage 1 of a gamma slot after its HowFar: SelectInteraction and the
// interaction-length update of Perform, or only that update on a boundary
// (G4HepEmGammaManager.icc:54-62,173-219; G4HepEmTrackingManager.cc:1092-1108).
template <typename R>
HD uint8_t GammaSelectAfterHowFar(const DevTables<R>& T, Trk<R>& t, R pe, Rng& rng, const hepem_tracks& in,
                                  hepem_tracks& out, const hepem_step_out& res, Staged* st, int n, int i) {
  uint8_t proc = kPNone;
  if (t.onBoundary) {
    t.nial[0] -= t.gstep / t.mfp[0];
  } else {
    SampleInteraction(T, t, pe, (R)rng.flat());
    if (t.winner != 3) {
      t.nial[0] -= t.gstep / t.mfp[0];
      ResetNumIALeft(t.nial, t.winner);
      proc = t.winner == 0 ? kPConv : t.winner == 1 ? kPCompton : kPPE;
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) out.num_ia_left[k * n + i] = (double)t.nial[k];
  res.true_step[i] = (double)t.gstep;
  res.geom_step[i] = (double)t.gstep;
#pragma unroll
  for (int c = 0; c < 3; ++c) res.displacement[c * n + i] = 0.0;
  res.winner[i] = t.winner;
  if (proc != kPNone) {
    const double d[3] = {(double)t.dir[0], (double)t.dir[1], (double)t.dir[2]};
    return StageTrack(st, i, proc, false, t.imc, in.id[i], StreamHi(in, i), t.ekin, d, rng, (double)pe);
  }
  out.ekin[i] = t.ekin;
#pragma unroll
  for (int c = 0; c < 3; ++c) out.dir[c * n + i] = (double)t.dir[c];
  StoreRng(out, i, rng);
  res.edep[i] = t.edep;
  res.stopped[i] = t.ekin <= 0.0 ? 1 : 0;
  SecList<R> sl;
  sl.ns = 0;
  WriteSecondaries(res, n, i, sl);
  return kPNone;
}

// Stage 1 of a gamma slot of the fused step: HowFar, then as above.
template <typename R>
HD uint8_t StepGammaSelect(const DevTables<R>& T, uint64_t seed, const hepem_tracks& in, hepem_tracks& out,
                           const hepem_step_out& res, Staged* st, int n, int i) {
  Trk<R> t;
  t.isEl = false;
  LoadTrack(in, n, i, t);
  Rng rng;
  rng.init(seed, (uint32_t)in.id[i], StreamHi(in, i), in.rng_ctr[i], in.rng_has_gauss[i] != 0, in.rng_gauss[i]);
  const R pe = GammaHowFar(T, t, rng);
  return GammaSelectAfterHowFar(T, t, pe, rng, in, out, res, st, n, i);
}

// Stage 1 of a gamma slot of the split Perform: step, mean free path and PE
// cross section from the hand-over.
template <typename R>
HD uint8_t PerformGammaSelect(const DevTables<R>& T, uint64_t seed, hepem_tracks& t, const hepem_howfar_state& hs,
                              const hepem_step_out& res, Staged* st, int n, int i) {
  Trk<R> g;
  g.isEl = false;
  LoadTrack(t, n, i, g);
  g.gstep = (R)hs.geom_step[i];
  g.mfp[0] = (R)hs.mfp[i];
  Rng rng;
  rng.init(seed, (uint32_t)t.id[i], StreamHi(t, i), t.rng_ctr[i], t.rng_has_gauss[i] != 0, t.rng_gauss[i]);
  return GammaSelectAfterHowFar(T, g, (R)hs.pe_xsec[i], rng, t, t, res, st, n, i);
}

// Stage 2: track state of a pending slot from its staged record.
template <typename R> HD void Unstage(const Staged& r, uint8_t code, uint64_t seed, Trk<R>& t, Rng& rng) {
  t.isEl = (code & kCodeElectron) != 0;
  t.setE(r.ekin);
#pragma unroll
  for (int c = 0; c < 3; ++c) t.dir[c] = (R)r.dir[c];
  t.imc = r.imc;
  t.edep = 0.0;
  rng.init(seed, (uint32_t)r.id, r.stream, r.ctr, (code & kCodeGauss) != 0, r.gauss);
}

// Stage 2 of an e-/e+ slot: the final state of the chosen interaction
// (PerformDiscrete, G4HepEmElectronManager.icc:443-458).
template <typename R>
HD void ElectronInteract(const DevTables<R>& T, uint64_t seed, hepem_tracks& out, const hepem_step_out& res, int n, int i,
                         const Staged& r, uint8_t code) {
  Trk<R> t;
  Rng rng;
  Unstage(r, code, seed, t, rng);
  SecList<R> sl;
  sl.ns = 0;
  switch (code & kCodeProc) {
    case kPIoniEl:
    case kPIoniPos: IoniPerform(T, t, sl, rng); break;
    case kPBremSB: BremPerform(T, t, sl, rng, true); break;
    case kPBremRB: BremPerform(T, t, sl, rng, false); break;
    default: AnnihilateInFlight(t, sl, rng); break;
  }
  out.ekin[i] = t.ekin;
#pragma unroll
  for (int c = 0; c < 3; ++c) out.dir[c * n + i] = (double)t.dir[c];
  StoreRng(out, i, rng);
  res.stopped[i] = t.ekin <= 0.0 ? 1 : 0;
  WriteSecondaries(res, n, i, sl);
}

// Stage 2 of a gamma slot: the interaction and the gamma tracking cut of
// G4HepEmGammaManager::Perform (G4HepEmGammaManager.icc:63-93).
template <typename R>
HD void GammaInteract(const DevTables<R>& T, uint64_t seed, hepem_tracks& out, const hepem_step_out& res, int n, int i,
                      const Staged& r, uint8_t code) {
  Trk<R> t;
  Rng rng;
  Unstage(r, code, seed, t, rng);
  SecList<R> sl;
  sl.ns = 0;
  switch (code & kCodeProc) {
    case kPConv: ConvPerform(T, t, sl, rng); break;
    case kPCompton: ComptonPerform(t, sl, rng); break;
    default: PEPerform(T, t, (R)r.pe, sl, rng); break;
  }
  const double finalEkin = t.ekin;
  if (finalEkin > 0.0 && finalEkin <= T.gamTrackCut) {
    t.setE(0.0);
    t.edep += finalEkin;
  }
  out.ekin[i] = t.ekin;
#pragma unroll
  for (int c = 0; c < 3; ++c) out.dir[c * n + i] = (double)t.dir[c];
  StoreRng(out, i, rng);
  res.edep[i] = t.edep;
  res.stopped[i] = t.ekin <= 0.0 ? 1 : 0;
  WriteSecondaries(res, n, i, sl);
}

// Process sort and final states, tile by tile (the single-batch two-stage
// step and Perform).
//
// k_classify_tiles lays the pending slots of each tile of kClassifyTile
// slots out process by process, within the tile's own range of `tlist` (tile
// offsets, 16 bits), and writes the tile's process offsets and chunk counts
// to `tinfo` (wave ballots and an LDS scan; no atomics, so the layout is
// deterministic).
//
// k_interact_tiles gives each wave one chunk of up to 64 entries of one
// process. A tile's chunks are spread over kInteractParts blocks that go to
// the same XCD one after another (blocks are dealt to the XCDs round-robin),
// so all final-state writes to the tile's cache lines come from one XCD
// within a short time and its L2 merges them into whole lines before they go
// to HBM. With a global process-sorted queue each wave's writes land a few
// slots per line and reach HBM as partial lines.
constexpr int kXcds = 8;
constexpr int kTileInfo = 2 * (kNumProc + 1);   // per tile: process offsets, first chunks
constexpr int kInteractWaves = kBlock / 64;
// blocks of a tile of `items` * kClassifyBlock slots: at most one chunk per 64
// slots plus one partial chunk per process, kInteractWaves chunks per block
constexpr int InteractParts(int items) { return (items * kClassifyBlock / 64 + kNumProc + kInteractWaves - 1) / kInteractWaves; }
constexpr int kInteractParts = InteractParts(kClassifyItems);

inline int InteractGridTiles(int tiles, int parts = kInteractParts) { return (tiles + kXcds - 1) / kXcds * kXcds * parts; }
inline int InteractGrid(int n) { return InteractGridTiles((n + kClassifyTile - 1) / kClassifyTile); }

// Tile-local particle partition of tile `tile` of a batch of n slots: the
// tile's e-/e+ slots first, then its gammas, each in slot order, at the
// tile's own range of the batch's queue; *ecount = the tile's e-/e+ count.
// One kClassifyBlock-thread block per tile.
__device__ void ClassifyLocalTile(int n, const int8_t* __restrict__ ptype, uint32_t* __restrict__ queue, int tile,
                                  uint32_t* __restrict__ ecount) {
  __shared__ uint32_t sE[kClassifyWaves], sG[kClassifyWaves], sTot;
  const int lane = __lane_id();
  const int wave = threadIdx.x >> 6;
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int tile0 = tile * kClassifyTile;
  uint32_t cE = 0, cG = 0;
  for (int it = 0; it < kClassifyItems; ++it) {
    const int i = tile0 + it * kClassifyBlock + threadIdx.x;
    const bool valid = i < n;
    const int8_t pt = valid ? ptype[i] : (int8_t)-1;
    const bool isE = valid && (pt == 0 || pt == 1);
    cE += (uint32_t)__popcll(__ballot(isE));
    cG += (uint32_t)__popcll(__ballot(valid && !isE));
  }
  if (lane == 0) {
    sE[wave] = cE;
    sG[wave] = cG;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tE = 0, tG = 0;
    for (int w = 0; w < kClassifyWaves; ++w) {
      const uint32_t e = sE[w], g = sG[w];
      sE[w] = tE;
      sG[w] = tG;
      tE += e;
      tG += g;
    }
    sTot = tE;
    *ecount = tE;
  }
  __syncthreads();
  uint32_t pE = sE[wave], pG = sTot + sG[wave];
  for (int it = 0; it < kClassifyItems; ++it) {
    const int i = tile0 + it * kClassifyBlock + threadIdx.x;
    const bool valid = i < n;
    const int8_t pt = valid ? ptype[i] : (int8_t)-1;
    const bool isE = valid && (pt == 0 || pt == 1);
    const bool isG = valid && !isE;
    const uint64_t mE = __ballot(isE), mG = __ballot(isG);
    if (isE) queue[tile0 + pE + __popcll(mE & below)] = (uint32_t)i;
    if (isG) queue[tile0 + pG + __popcll(mG & below)] = (uint32_t)i;
    pE += (uint32_t)__popcll(mE);
    pG += (uint32_t)__popcll(mG);
  }
}

// The tile-local particle partition of one batch, one block per tile, with
// tinfo[tile * kTileInfo] = the tile's e-/e+ count. The first block also
// clears the step's counters and marks the layout in counters[kCounterLayout]
// (no counter is read before the next kernel).
__global__ void __launch_bounds__(kClassifyBlock) k_classify_local(int n, const int8_t* __restrict__ ptype,
                                                                   uint32_t* __restrict__ queue, uint32_t* __restrict__ tinfo,
                                                                   uint32_t* __restrict__ counters) {
  if (blockIdx.x == 0 && threadIdx.x < kCounterWords) counters[threadIdx.x] = threadIdx.x == kCounterLayout ? 1u : 0u;
  ClassifyLocalTile(n, ptype, queue, (int)blockIdx.x, tinfo + (size_t)blockIdx.x * kTileInfo);
}

constexpr int kTileChunks = kClassifyTile / kBlock;
// tiles at the end of each XCD's share whose second halves (mostly gammas,
// short) run after all their first halves (mostly e-/e+, long)
constexpr int kTailTiles = 8;
static_assert(kTileChunks % 2 == 0, "a tile splits into two halves of whole blocks");
inline int LocalSelectGridTiles(int ntiles) { return (ntiles + kXcds - 1) / kXcds * kXcds * kTileChunks; }
inline int LocalSelectGrid(int n) { return LocalSelectGridTiles((n + kClassifyTile - 1) / kClassifyTile); }

// Block -> (tile, chunk of kBlock slots) of stage 1 over the tile-local
// particle partition of ntiles tiles: XCD x takes tiles x, x + kXcds, ...,
// each as kTileChunks consecutive blocks of its share, so a tile's e-/e+ and
// gamma blocks run on one XCD close together and its lines are read and
// written through one L2. In the last kTailTiles tiles of a share the blocks
// of the first halves come first, then those of the second halves, so the
// launch ends on short gamma blocks rather than on long e-/e+ ones. False for
// the blocks of a grid of LocalSelectGridTiles(ntiles) that have no chunk.
__device__ inline bool LocalSelectChunk(int ntiles, int& tile, int& chunk) {
  const int x = blockIdx.x % kXcds, q = blockIdx.x / kXcds;
  const int share = ntiles > x ? (ntiles - x + kXcds - 1) / kXcds : 0;
  const int tail = min(kTailTiles, share), body = (share - tail) * kTileChunks;
  int k;
  if (q < body) {
    k = q / kTileChunks;
    chunk = q % kTileChunks;
  } else {
    const int r = q - body, h = kTileChunks / 2;
    if (r >= tail * kTileChunks) return false;
    const bool first = r < tail * h;
    const int r2 = first ? r : r - tail * h;
    k = share - tail + r2 / h;
    chunk = (first ? 0 : h) + r2 % h;
  }
  tile = k * kXcds + x;
  return true;
}

// Stage 1 of the step over one batch's particle queue, its blocks placed tile by
// tile as LocalSelectChunk places them (the tile-local partition that
// hepem_step_prepare lays out with the process sort; a queue laid out over the
// whole batch is read correctly too).
template <typename R>
__global__ void __launch_bounds__(kBlock) k_step_select_local(DevTables<R> T, uint64_t seed, hepem_tracks in, hepem_tracks out,
                                                        hepem_step_out res, const uint32_t* __restrict__ queue,
                                                        uint32_t* __restrict__ counters, const uint32_t* __restrict__ tinfo,
                                                        uint8_t* __restrict__ code, Staged* __restrict__ st) {
  const int n = in.n;
  int tile, chunk;
  if (!LocalSelectChunk((n + kClassifyTile - 1) / kClassifyTile, tile, chunk)) return;
  const int tile0 = tile * kClassifyTile;
  const int m = min(kClassifyTile, n - tile0);
  const int jl = chunk * kBlock + (int)threadIdx.x;
  if (jl >= m) return;
  // the layout hepem_step_prepare chose: tile-local, or e-/e+ first over the batch
  const bool isE = counters[kCounterLayout] ? jl < (int)tinfo[(size_t)tile * kTileInfo] : (uint32_t)(tile0 + jl) < counters[0];
  const int i = (int)queue[tile0 + jl];
  if (RejectTrack(T, in, out, res, n, i, counters)) {
    code[i] = kPNone;
    return;
  }
  CopyIdentity(in, out, i);
  code[i] = isE ? StepElectronSelect(T, seed, in, out, res, st, n, i) : StepGammaSelect(T, seed, in, out, res, st, n, i);
}

// Thread -> (slot, is-e-/e+) of the split HowFar and Perform kernels. A
// mixed batch's queue is read in the layout its hepem_howfar prepare chose
// (counters[kCounterLayout]): tile-local, with the blocks placed as
// k_step_select_local places them (grid LocalSelectGrid(n)), or e-/e+ first
// over the batch (one thread per entry). False if the thread has no slot;
// kindOk false if the slot holds a track of the other kind.
template <int KIND>
__device__ inline bool ThreadSlot(const hepem_tracks& t, const uint32_t* __restrict__ queue,
                                  const uint32_t* __restrict__ counters, const uint32_t* __restrict__ tinfo, int& i,
                                  bool& isE, bool& kindOk) {
  const int n = t.n;
  if (KIND == kMixed && counters[kCounterLayout]) {
    int tile, chunk;
    if (!LocalSelectChunk((n + kClassifyTile - 1) / kClassifyTile, tile, chunk)) return false;
    const int tile0 = tile * kClassifyTile, jl = chunk * kBlock + (int)threadIdx.x;
    if (tile0 + jl >= n) return false;
    i = (int)queue[tile0 + jl];
    isE = jl < (int)tinfo[(size_t)tile * kTileInfo];
    kindOk = true;
    return true;
  }
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return false;
  kindOk = ResolveSlot<KIND>(t, j, queue, counters, i, isE);
  return true;
}

template <typename R, int KIND>
__global__ void __launch_bounds__(kBlock, 4) k_howfar(DevTables<R> T, uint64_t seed, hepem_tracks t, hepem_howfar_state hs,
                                                   const uint32_t* __restrict__ queue, uint32_t* __restrict__ counters,
                                                   const uint32_t* __restrict__ tinfo) {
  const int n = t.n;
  int i;
  bool isE, kindOk;
  if (!ThreadSlot<KIND>(t, queue, counters, tinfo, i, isE, kindOk)) return;
  if (!kindOk || !TrackValid(T, t, i)) {
    atomicAdd(&counters[2], 1u);
    hs.winner[i] = -99;
    hs.geom_step[i] = 0.0;
    return;
  }
  if (isE) {
    ElState<R> el;
    Rng rng;
    InitElectron(t, n, i, seed, el, rng);
    ElectronHowFar(T, el, rng);
    StoreElectronState(t, n, i, el, rng);
    StoreElectronHowFar(hs, n, i, el);
  } else {
    Trk<R> g;
    g.isEl = false;
    LoadTrack(t, n, i, g);
    Rng rng;
    rng.init(seed, (uint32_t)t.id[i], StreamHi(t, i), t.rng_ctr[i], t.rng_has_gauss[i] != 0, t.rng_gauss[i]);
    const R pe = GammaHowFar(T, g, rng);
#pragma unroll
    for (int k = 0; k < 4; ++k) t.num_ia_left[k * n + i] = (double)g.nial[k];
    StoreRng(t, i, rng);
    hs.geom_step[i] = (double)g.gstep;
    hs.mfp[i] = (double)g.mfp[0];
    hs.winner[i] = g.winner;
    hs.pe_xsec[i] = (double)pe;
  }
}

template <typename R, int KIND>
__global__ void __launch_bounds__(kBlock, 3) k_perform(DevTables<R> T, uint64_t seed, hepem_tracks t, hepem_howfar_state hs,
                                                    hepem_step_out res, const uint32_t* __restrict__ queue,
                                                    uint32_t* __restrict__ counters, const uint32_t* __restrict__ tinfo) {
  const int n = t.n;
  int i;
  bool isE, kindOk;
  if (!ThreadSlot<KIND>(t, queue, counters, tinfo, i, isE, kindOk)) return;
  if (!kindOk || hs.winner[i] == -99) {
    WriteRejected<R>(t, t, res, n, i);
    return;
  }
  SecList<R> sl;
  sl.ns = 0;
  if (isE) {
    ElState<R> el;
    Rng rng;
    InitElectron(t, n, i, seed, el, rng);
    LoadElectronHowFar(hs, n, i, el);
    ElectronPerform(T, el, sl, rng);
    StoreElectronState(t, n, i, el, rng);
    StoreElectronResult(t, res, n, i, el, sl);
  } else {
    Trk<R> g;
    g.isEl = false;
    LoadTrack(t, n, i, g);
    g.gstep = (R)hs.geom_step[i];
    g.mfp[0] = (R)hs.mfp[i];
    Rng rng;
    rng.init(seed, (uint32_t)t.id[i], StreamHi(t, i), t.rng_ctr[i], t.rng_has_gauss[i] != 0, t.rng_gauss[i]);
    GammaPerform(T, g, (R)hs.pe_xsec[i], sl, rng);
    StoreGamma(t, res, n, i, g, sl, rng);
  }
}

// Stage 1 of the split Perform (as k_perform up to the choice of the
// interaction).
template <typename R, int KIND>
__global__ void __launch_bounds__(kBlock) k_perform_select(DevTables<R> T, uint64_t seed, hepem_tracks t,
                                                           hepem_howfar_state hs, hepem_step_out res,
                                                           const uint32_t* __restrict__ queue, uint32_t* __restrict__ counters,
                                                           const uint32_t* __restrict__ tinfo, uint8_t* __restrict__ code,
                                                           Staged* __restrict__ st) {
  const int n = t.n;
  int i;
  bool isE, kindOk;
  if (!ThreadSlot<KIND>(t, queue, counters, tinfo, i, isE, kindOk)) return;
  if (!kindOk || hs.winner[i] == -99) {
    WriteRejected<R>(t, t, res, n, i);
    code[i] = kPNone;
    return;
  }
  code[i] = isE ? PerformElectronSelect<R>(T, seed, t, hs, res, st, n, i) : PerformGammaSelect<R>(T, seed, t, hs, res, st, n, i);
}

// The fused step over either particle layout of the queue (the one
// hepem_step_prepare chose, as counters[kCounterLayout] records): e-/e+ are
// the first counters[0] entries, or the first tinfo[tile * kTileInfo] entries
// of each tile.
template <typename R>
__global__ void __launch_bounds__(kBlock) k_step(DevTables<R> T, uint64_t seed, hepem_tracks in, hepem_tracks out,
                                                 hepem_step_out res, const uint32_t* __restrict__ queue,
                                                 uint32_t* __restrict__ counters, const uint32_t* __restrict__ tinfo) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = in.n;
  if (j >= n) return;
  const bool isE = counters[kCounterLayout] ? j % kClassifyTile < (int)tinfo[(size_t)(j / kClassifyTile) * kTileInfo]
                                            : (uint32_t)j < counters[0];
  const int i = (int)queue[j];
  if (RejectTrack(T, in, out, res, n, i, counters)) return;
  CopyIdentity(in, out, i);
  if (isE) StepElectron(T, seed, in, out, res, n, i);
  else StepGamma(T, seed, in, out, res, n, i);
}

// the process layout of tile `tile` of a batch of n slots (codes, tile list
// and the tile's info words of that batch); a tile is ITEMS * kClassifyBlock
// slots
template <int ITEMS = kClassifyItems>
__device__ void ClassifyTileLayout(int n, const uint8_t* __restrict__ code, uint16_t* __restrict__ tlist,
                                   uint32_t* __restrict__ info, int tile) {
  static_assert(ITEMS * kClassifyBlock <= 65536, "tile-local offsets are 16 bits");
  __shared__ uint32_t sW[kNumProc][kClassifyWaves];
  const int lane = __lane_id();
  const int wave = threadIdx.x >> 6;
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int tile0 = tile * ITEMS * kClassifyBlock;
  uint32_t cnt[kNumProc];
#pragma unroll
  for (int q = 0; q < kNumProc; ++q) cnt[q] = 0u;
  for (int it = 0; it < ITEMS; ++it) {
    const int i = tile0 + it * kClassifyBlock + threadIdx.x;
    const int c = i < n ? (int)(code[i] & kCodeProc) : 0;
#pragma unroll
    for (int q = 0; q < kNumProc; ++q) cnt[q] += (uint32_t)__popcll(__ballot(c == q + 1));
  }
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < kNumProc; ++q) sW[q][wave] = cnt[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t off = 0u, chunk = 0u;
    for (int q = 0; q < kNumProc; ++q) {
      info[q] = off;
      info[kNumProc + 1 + q] = chunk;
      const uint32_t first = off;
      for (int w = 0; w < kClassifyWaves; ++w) {
        const uint32_t v = sW[q][w];
        sW[q][w] = off;
        off += v;
      }
      chunk += (off - first + 63u) / 64u;
    }
    info[kNumProc] = off;
    info[2 * kNumProc + 1] = chunk;
  }
  __syncthreads();
  uint32_t pos[kNumProc];
#pragma unroll
  for (int q = 0; q < kNumProc; ++q) pos[q] = sW[q][wave];
  for (int it = 0; it < ITEMS; ++it) {
    const int local = it * kClassifyBlock + threadIdx.x;
    const int i = tile0 + local;
    const int c = i < n ? (int)(code[i] & kCodeProc) : 0;
#pragma unroll
    for (int q = 0; q < kNumProc; ++q) {
      const uint64_t m = __ballot(c == q + 1);
      if (c == q + 1) tlist[tile0 + pos[q] + __popcll(m & below)] = (uint16_t)local;
      pos[q] += (uint32_t)__popcll(m);
    }
  }
}

__global__ void __launch_bounds__(kClassifyBlock) k_classify_tiles(int n, const uint8_t* __restrict__ code,
                                                                   uint16_t* __restrict__ tlist, uint32_t* __restrict__ tinfo) {
  ClassifyTileLayout(n, code, tlist, tinfo + (size_t)blockIdx.x * kTileInfo, (int)blockIdx.x);
}

// block -> (tile, chunk of this wave): XCD x takes tiles x, x + kXcds, ...,
// each as kInteractParts consecutive blocks of its share; false past the last
// tile
template <int PARTS = kInteractParts> __device__ inline bool TileChunk(uint32_t ntiles, uint32_t& tile, uint32_t& chunk) {
  const uint32_t x = blockIdx.x % kXcds;
  const uint32_t inXcd = blockIdx.x / kXcds;
  tile = (inXcd / PARTS) * kXcds + x;
  chunk = (inXcd % PARTS) * kInteractWaves + (threadIdx.x >> 6);
  return tile < ntiles;
}

// final state of this lane's entry of chunk `chunk` of a tile whose first
// slot is tile0 (codes, staged records and tile list of its batch)
template <typename R>
__device__ void InteractTileChunk(const DevTables<R>& T, uint64_t seed, hepem_tracks& out, const hepem_step_out& res, int n,
                                  const uint8_t* __restrict__ code, const Staged* __restrict__ st,
                                  const uint16_t* __restrict__ tlist, const uint32_t* __restrict__ info, int tile0,
                                  uint32_t chunk) {
  if (chunk >= info[2 * kNumProc + 1]) return;
  int q = 0;
  while (q + 1 < kNumProc && chunk >= info[kNumProc + 2 + q]) ++q;
  const uint32_t k = (chunk - info[kNumProc + 1 + q]) * 64u + (uint32_t)__lane_id();
  if (k >= info[q + 1] - info[q]) return;
  const int i = tile0 + (int)tlist[tile0 + info[q] + k];
  const uint8_t c = code[i];
  const Staged r = st[i];
  if (q + 1 < kPConv) ElectronInteract<R>(T, seed, out, res, n, i, r, c);
  else GammaInteract<R>(T, seed, out, res, n, i, r, c);
}

template <typename R>
__global__ void __launch_bounds__(kBlock) k_interact_tiles(DevTables<R> T, uint64_t seed, hepem_tracks out, hepem_step_out res,
                                                           const uint8_t* __restrict__ code, const Staged* __restrict__ st,
                                                           const uint16_t* __restrict__ tlist,
                                                           const uint32_t* __restrict__ tinfo) {
  const int n = out.n;
  uint32_t tile, chunk;
  if (!TileChunk((uint32_t)((n + kClassifyTile - 1) / kClassifyTile), tile, chunk)) return;
  InteractTileChunk<R>(T, seed, out, res, n, code, st, tlist, tinfo + (size_t)tile * kTileInfo, (int)tile * kClassifyTile,
                       chunk);
}

// ---- batched stepping ---------------------------------------------------------------------------
// Block sizes of the batched step kernel, selected at run time.
constexpr int kNumShapes = 3;
constexpr int ShapeBlock(int s) { return 64 << s; }
// Process-sort tile sizes of the batched step, selected at run time:
// TileSlots(t) = 1024 << t slots (1024 ... 8192), TileItems(t) per thread of
// the classification block.
constexpr int kNumTiles = 4;
constexpr int TileItems(int t) { return 1 << t; }
constexpr int TileSlots(int t) { return kClassifyBlock << t; }
// the tile size of the particle classification among them
constexpr int kParticleTiles = 2;
static_assert(TileSlots(kParticleTiles) == kClassifyTile, "particle tiles are one of the sort tile sizes");
// Words of a plan's work counters after its per-batch counters: the next
// chunk of the persistent step kernel, then the next work item of each XCD's
// share of the persistent tile kernel.
constexpr int kQueueWords = 1 + kXcds;

// Device descriptor of one batch of a plan. A plan holds nbatch + 1 of them;
// the last carries only the totals of the prefix fields.
struct BatchDev {
  hepem_tracks in, out;
  hepem_step_out res;
  uint32_t qoff;                 // first entry of the batch's queue
  uint32_t tile0;                // first particle-classify tile of the batch (kClassifyTile slots)
  uint32_t ptile0[kNumTiles];    // first process-sort tile of the batch for tile size TileSlots(t)
  uint32_t chunk0[kNumShapes];   // first step chunk for block size ShapeBlock(s)
};


// bcounters[4b + 0..2]: e-/e+, gammas, rejected tracks of batch b
__global__ void __launch_bounds__(kClassifyBlock) k_classify_batched(const BatchDev* __restrict__ d,
                                                             const uint32_t* __restrict__ tbatch,
                                                             uint32_t* __restrict__ queue, uint32_t* __restrict__ bcounters) {
  const uint32_t t = blockIdx.x;
  const int b = __builtin_amdgcn_readfirstlane((int)tbatch[t]);
  ClassifyTile(d[b].in.n, d[b].in.ptype, queue + d[b].qoff, bcounters + 4 * b, (int)(t - d[b].tile0));
}

// One thread per queue entry of every batch: block c takes chunk c of the
// chunk sequence (all batches' chunks of BS entries, in batch order). The
// batch index is uniform over the block: it (and the descriptor loads) stay
// scalar.
template <typename R, int S>
__device__ void StepChunk(const DevTables<R>& T, uint64_t seed, const BatchDev* __restrict__ d,
                          const uint32_t* __restrict__ cbatch, uint32_t c,
                          const uint32_t* __restrict__ queue, uint32_t* __restrict__ bcounters) {
  constexpr int BS = ShapeBlock(S);
  const int b = __builtin_amdgcn_readfirstlane((int)cbatch[c]);
  const hepem_tracks in = d[b].in;
  hepem_tracks out = d[b].out;
  const hepem_step_out res = d[b].res;
  const int n = in.n;
  const int j = (int)(c - d[b].chunk0[S]) * BS + (int)threadIdx.x;
  if (j >= n) return;
  uint32_t* cnt = bcounters + 4 * b;
  const int i = (int)queue[d[b].qoff + j];
  if (RejectTrack(T, in, out, res, n, i, cnt)) return;
  CopyIdentity(in, out, i);
  if ((uint32_t)j < cnt[0]) StepElectron(T, seed, in, out, res, n, i);
  else StepGamma(T, seed, in, out, res, n, i);
}

// Stage 1 of the two-stage form over every batch: records the process code
// and staged record of slot i of batch b at global index qoff_b + i;
// k_classify_tiles_batched and k_interact_tiles_batched then sort and sample
// them tile by tile, as the single-batch step does (tiles numbered across
// the batches).
template <typename R, int S>
__device__ void StepSelectChunk(const DevTables<R>& T, uint64_t seed, const BatchDev* __restrict__ d,
                                const uint32_t* __restrict__ cbatch, uint32_t c,
                                const uint32_t* __restrict__ queue, uint32_t* __restrict__ bcounters,
                                uint8_t* __restrict__ code, Staged* __restrict__ st) {
  constexpr int BS = ShapeBlock(S);
  const int b = __builtin_amdgcn_readfirstlane((int)cbatch[c]);
  const hepem_tracks in = d[b].in;
  hepem_tracks out = d[b].out;
  const hepem_step_out res = d[b].res;
  const int n = in.n;
  const int j = (int)(c - d[b].chunk0[S]) * BS + (int)threadIdx.x;
  if (j >= n) return;
  uint32_t* cnt = bcounters + 4 * b;
  const uint32_t qoff = d[b].qoff;
  const int i = (int)queue[qoff + j];
  uint8_t* cb = code + qoff;
  if (RejectTrack(T, in, out, res, n, i, cnt)) {
    cb[i] = kPNone;
    return;
  }
  CopyIdentity(in, out, i);
  cb[i] = (uint32_t)j < cnt[0] ? StepElectronSelect(T, seed, in, out, res, st + qoff, n, i)
                               : StepGammaSelect(T, seed, in, out, res, st + qoff, n, i);
}

// The tile-local particle partition over every batch: block t lays out
// particle-classify tile t of the plan (tiles numbered across the batches, as
// BatchDev::tile0), within its batch's queue; ecount[t] = its e-/e+ count.
__global__ void __launch_bounds__(kClassifyBlock) k_classify_local_batched(const BatchDev* __restrict__ d,
                                                                           const uint32_t* __restrict__ tbatch,
                                                                           uint32_t* __restrict__ queue,
                                                                           uint32_t* __restrict__ ecount) {
  const uint32_t t = blockIdx.x;
  const int b = __builtin_amdgcn_readfirstlane((int)tbatch[t]);
  ClassifyLocalTile(d[b].in.n, d[b].in.ptype, queue + d[b].qoff, (int)(t - d[b].tile0), ecount + t);
}

// Stage 1 of the two-stage form over the tile-local partition of every
// batch (kBlock threads, a grid of LocalSelectGridTiles(ntiles)): the blocks
// of the plan's tiles are placed as k_step_select_local places a single
// batch's, and each records what StepSelectChunk records.
template <typename R>
__global__ void __launch_bounds__(kBlock) k_step_select_local_batched(DevTables<R> T, uint64_t seed,
                                                                      const BatchDev* __restrict__ d,
                                                                      const uint32_t* __restrict__ tbatch, int ntiles,
                                                                      const uint32_t* __restrict__ queue,
                                                                      uint32_t* __restrict__ bcounters,
                                                                      const uint32_t* __restrict__ ecount,
                                                                      uint8_t* __restrict__ code, Staged* __restrict__ st) {
  int tile, chunk;
  if (!LocalSelectChunk(ntiles, tile, chunk)) return;
  const int b = __builtin_amdgcn_readfirstlane((int)tbatch[tile]);
  const hepem_tracks in = d[b].in;
  hepem_tracks out = d[b].out;
  const hepem_step_out res = d[b].res;
  const int n = in.n;
  const int tile0 = (tile - (int)d[b].tile0) * kClassifyTile;
  const int m = min(kClassifyTile, n - tile0);
  const int jl = chunk * kBlock + (int)threadIdx.x;
  if (jl >= m) return;
  const uint32_t qoff = d[b].qoff;
  const int i = (int)queue[qoff + tile0 + jl];
  uint8_t* cb = code + qoff;
  if (RejectTrack(T, in, out, res, n, i, bcounters + 4 * b)) {
    cb[i] = kPNone;
    return;
  }
  CopyIdentity(in, out, i);
  cb[i] = jl < (int)ecount[tile] ? StepElectronSelect(T, seed, in, out, res, st + qoff, n, i)
                                 : StepGammaSelect(T, seed, in, out, res, st + qoff, n, i);
}

// Next work item of a persistent kernel: thread 0 takes it from *next and the
// block reads it from LDS (the second barrier keeps thread 0 from overwriting
// it before every thread has read it).
__device__ inline uint32_t NextItem(uint32_t* next) {
  __shared__ uint32_t sItem;
  if (threadIdx.x == 0) sItem = atomicAdd(next, 1u);
  __syncthreads();
  const uint32_t c = sItem;
  __syncthreads();
  return c;
}

// Dispatched form: block c takes chunk c.
template <typename R, int S>
__global__ void __launch_bounds__(ShapeBlock(S)) k_step_batched(DevTables<R> T, uint64_t seed, const BatchDev* __restrict__ d,
                                                                const uint32_t* __restrict__ cbatch, uint32_t nchunks,
                                                                const uint32_t* __restrict__ queue,
                                                                uint32_t* __restrict__ bcounters) {
  if (blockIdx.x < nchunks) StepChunk<R, S>(T, seed, d, cbatch, blockIdx.x, queue, bcounters);
}

// Persistent form: one resident round of blocks, each taking chunks from the
// counter *next (zero at launch) until they run out. Nothing but the chunk
// index is carried from one chunk to the next.
template <typename R, int S>
__global__ void __launch_bounds__(ShapeBlock(S)) k_step_batched_q(DevTables<R> T, uint64_t seed, const BatchDev* __restrict__ d,
                                                                  const uint32_t* __restrict__ cbatch, uint32_t nchunks,
                                                                  const uint32_t* __restrict__ queue,
                                                                  uint32_t* __restrict__ bcounters, uint32_t* __restrict__ next) {
  for (uint32_t c = NextItem(next); c < nchunks; c = NextItem(next)) StepChunk<R, S>(T, seed, d, cbatch, c, queue, bcounters);
}

template <typename R, int S>
__global__ void __launch_bounds__(ShapeBlock(S)) k_step_select_batched(DevTables<R> T, uint64_t seed,
                                                                       const BatchDev* __restrict__ d, const uint32_t* __restrict__ cbatch,
                                                                       uint32_t nchunks, const uint32_t* __restrict__ queue,
                                                                       uint32_t* __restrict__ bcounters,
                                                                       uint8_t* __restrict__ code, Staged* __restrict__ st) {
  if (blockIdx.x < nchunks) StepSelectChunk<R, S>(T, seed, d, cbatch, blockIdx.x, queue, bcounters, code, st);
}

template <typename R, int S>
__global__ void __launch_bounds__(ShapeBlock(S)) k_step_select_batched_q(DevTables<R> T, uint64_t seed,
                                                                         const BatchDev* __restrict__ d, const uint32_t* __restrict__ cbatch,
                                                                         uint32_t nchunks, const uint32_t* __restrict__ queue,
                                                                         uint32_t* __restrict__ bcounters,
                                                                         uint8_t* __restrict__ code, Staged* __restrict__ st,
                                                                         uint32_t* __restrict__ next) {
  for (uint32_t c = NextItem(next); c < nchunks; c = NextItem(next))
    StepSelectChunk<R, S>(T, seed, d, cbatch, c, queue, bcounters, code, st);
}

// The tile-local process sort over all batches of a plan, in tiles of
// TileSlots(TI) slots: tiles are numbered across the batches
// (BatchDev::ptile0[TI]), each batch's tile list and codes sit at its queue
// offset.
template <int TI>
__global__ void __launch_bounds__(kClassifyBlock) k_classify_tiles_batched(const BatchDev* __restrict__ d,
                                                                           const uint32_t* __restrict__ tbatch,
                                                                           const uint8_t* __restrict__ code,
                                                                           uint16_t* __restrict__ tlist,
                                                                           uint32_t* __restrict__ tinfo) {
  const uint32_t t = blockIdx.x;
  const int b = __builtin_amdgcn_readfirstlane((int)tbatch[t]);
  const uint32_t qoff = d[b].qoff;
  ClassifyTileLayout<TileItems(TI)>(d[b].in.n, code + qoff, tlist + qoff, tinfo + (size_t)t * kTileInfo,
                                    (int)(t - d[b].ptile0[TI]));
}

template <typename R, int TI>
__device__ void InteractBatchedTile(const DevTables<R>& T, uint64_t seed, const BatchDev* __restrict__ d,
                                    const uint32_t* __restrict__ tbatch, uint32_t tile, uint32_t chunk,
                                    const uint8_t* __restrict__ code, const Staged* __restrict__ st,
                                    const uint16_t* __restrict__ tlist, const uint32_t* __restrict__ tinfo) {
  const int b = __builtin_amdgcn_readfirstlane((int)tbatch[tile]);
  hepem_tracks out = d[b].out;
  const uint32_t qoff = d[b].qoff;
  InteractTileChunk<R>(T, seed, out, d[b].res, out.n, code + qoff, st + qoff, tlist + qoff, tinfo + (size_t)tile * kTileInfo,
                       (int)(tile - d[b].ptile0[TI]) * TileSlots(TI), chunk);
}

template <typename R, int TI>
__global__ void __launch_bounds__(kBlock) k_interact_tiles_batched(DevTables<R> T, uint64_t seed, const BatchDev* __restrict__ d,
                                                                   const uint32_t* __restrict__ tbatch, uint32_t ntiles,
                                                                   const uint8_t* __restrict__ code,
                                                                   const Staged* __restrict__ st,
                                                                   const uint16_t* __restrict__ tlist,
                                                                   const uint32_t* __restrict__ tinfo) {
  uint32_t tile, chunk;
  if (!TileChunk<InteractParts(TileItems(TI))>(ntiles, tile, chunk)) return;
  InteractBatchedTile<R, TI>(T, seed, d, tbatch, tile, chunk, code, st, tlist, tinfo);
}

// Persistent form of k_interact_tiles_batched: one resident round of blocks.
// Block b works for XCD b % kXcds (blocks are dealt to the XCDs round-robin)
// and takes that XCD's work items, tiles x, x + kXcds, ... in
// InteractParts blocks' worth each, from its counter next[x] (zero at launch).
template <typename R, int TI>
__global__ void __launch_bounds__(kBlock) k_interact_tiles_batched_q(DevTables<R> T, uint64_t seed,
                                                                     const BatchDev* __restrict__ d,
                                                                     const uint32_t* __restrict__ tbatch, uint32_t ntiles,
                                                                     const uint8_t* __restrict__ code,
                                                                     const Staged* __restrict__ st,
                                                                     const uint16_t* __restrict__ tlist,
                                                                     const uint32_t* __restrict__ tinfo,
                                                                     uint32_t* __restrict__ next) {
  constexpr uint32_t P = InteractParts(TileItems(TI));
  const uint32_t x = blockIdx.x % kXcds;
  for (uint32_t k = NextItem(next + x);; k = NextItem(next + x)) {
    const uint32_t tile = (k / P) * kXcds + x;
    if (tile >= ntiles) return;
    const uint32_t chunk = (k % P) * kInteractWaves + (threadIdx.x >> 6);
    InteractBatchedTile<R, TI>(T, seed, d, tbatch, tile, chunk, code, st, tlist, tinfo);
  }
}

// the five config-1 quantities of query i at (mc, E, lnE): restricted dE/dx,
// range, inverse range of that range, ioni and brem cross sections; NaN for
// an invalid couple. Those are computed at couple 0 and then replaced, so
// every lane stores at the same places in the code (no branch around the
// memory operations: the wait for the next query's inputs then covers no
// store that a branch left unaccounted). The results are streamed out
// (non-temporal stores): nothing in the kernel reads them back.
template <typename R>
HD void LookupStore(const DevEl<R>& e, int nmc, int n, int i, int mc, double E, double lnE, double* __restrict__ out) {
  const size_t sn = (size_t)n;  // 5 n entries may exceed the int range
  const bool ok = mc >= 0 && mc < nmc;
  const double nan = __longlong_as_double(0x7FF8000000000000LL);
  mc = ok ? mc : 0;
  const R lE = (R)lnE;
  // the range's loss-grid bin is where its inverse is searched first (the
  // spline's grid energies are the inverse's ordinates there)
  const int bin = LogBin(lE, e.lossLogMin, e.lossILD, e.N);
  const R r = GetRestRange(e, mc, E, lE);
  __builtin_nontemporal_store(ok ? (double)GetRestDEDX(e, mc, E, lE) : nan, out + i);
  __builtin_nontemporal_store(ok ? (double)r : nan, out + sn + i);
  __builtin_nontemporal_store(ok ? GetInvRangeInBin(e, mc, (double)r, bin, e.lossGrid64[bin], e.lossGrid64[bin + 1]) : nan, out + 2 * sn + i);
  __builtin_nontemporal_store(ok ? (double)GetRestMacXSec(e, mc, E, lE, true) : nan, out + 3 * sn + i);
  __builtin_nontemporal_store(ok ? (double)GetRestMacXSec(e, mc, E, lE, false) : nan, out + 4 * sn + i);
}

template <typename R>
__global__ void __launch_bounds__(kBlock) k_lookup(DevTables<R> T, int ptype, int n, const int32_t* __restrict__ imc,
                                                   const double* __restrict__ ekin, const double* __restrict__ lekin,
                                                   double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  LookupStore(T.el[ptype], T.nmc, n, i, imc[i], ekin[i], lekin[i], out);
}

// Gamma macroscopic cross sections and the interaction they select, one query
// per thread, as the reference's TestMacXSecDataKernel drives
// G4HepEmGammaManager::GetTotalMacXSec and SampleInteraction
// (testing/GammaXSections/src/GammaMacXSecs.cu:20-31): the total cross section,
// the photoelectric one the track holds after SampleInteraction, and the
// process chosen for the uniform draw u. An invalid couple gives NaN and -99.
template <typename R>
__global__ void __launch_bounds__(kBlock) k_gamma_lookup(DevTables<R> T, int n, const int32_t* __restrict__ imc,
                                                         const double* __restrict__ ekin, const double* __restrict__ lekin,
                                                         const double* __restrict__ urnd, double* __restrict__ out_xs,
                                                         double* __restrict__ out_pe, int32_t* __restrict__ out_pid) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int mc = imc[i];
  if (mc < 0 || mc >= T.nmc) {
    out_xs[i] = out_pe[i] = __longlong_as_double(0x7FF8000000000000LL);
    out_pid[i] = -99;
    return;
  }
  Trk<R> t;
  t.ekin = ekin[i];
  t.lekin = (R)lekin[i];
  t.imc = mc;
  t.winner = -1;
  for (int k = 0; k < 4; ++k) t.mfp[k] = t.nial[k] = (R)0.0;
  R pe = (R)0.0;
  const R tot = GetTotalMacXSec(T, t, pe);
  t.mfp[0] = (tot > (R)0.) ? (R)1. / tot : (R)kALargeValue;
  SampleInteraction(T, t, pe, (R)urnd[i]);
  out_xs[i] = (double)tot;
  out_pe[i] = (double)pe;
  out_pid[i] = t.winner;
}

// LDS staging of the config-1 lookups. The loss grid and data and the
// cross-section blocks of every couple (the doubles, and in the FP32 form
// the floats too) are a few kB per couple; each block copies them to LDS
// once and then serves a contiguous chunk of the queries from there, so the
// table gathers leave the texture path to the query stream.
constexpr size_t kLookupLdsMax = 64 * 1024;  // larger images use k_lookup
constexpr int kLookupBlocks = 2048;          // resident blocks if the occupancy query gives nothing
constexpr int kLookupPerThread = 16;         // queries per thread of a chunk, at most

__host__ __device__ inline size_t Al16(size_t b) { return (b + 15) & ~size_t(15); }

// LDS bytes of the image (the arrays in the order k_lookup_lds stages them)
template <typename R> __host__ __device__ size_t LookupLdsBytes(const DevEl<R>& e, int nmc) {
  const size_t N = (size_t)e.N, L = 5 * N * (size_t)nmc, X = (size_t)e.xsN;
  size_t b = Al16(N * sizeof(double)) + Al16(L * sizeof(double)) + Al16(X * sizeof(double)) +
             Al16(2 * (size_t)nmc * sizeof(XsBlock<R>));
  if (sizeof(R) != sizeof(double)) b += Al16(N * sizeof(R)) + Al16(L * sizeof(R)) + Al16(X * sizeof(R));
  return b;
}

// copies `bytes` (a multiple of 4) from src to lds + off with the whole block;
// returns the LDS copy and advances off
__device__ inline const void* StageArray(unsigned char* lds, size_t& off, const void* src, size_t bytes) {
  unsigned char* dst = lds + off;
  if ((bytes & 15) == 0) {
    const uint4* s4 = reinterpret_cast<const uint4*>(src);
    uint4* d4 = reinterpret_cast<uint4*>(dst);
    for (size_t k = threadIdx.x; k < bytes / 16; k += blockDim.x) d4[k] = s4[k];
  } else {
    const uint32_t* s1 = reinterpret_cast<const uint32_t*>(src);
    uint32_t* d1 = reinterpret_cast<uint32_t*>(dst);
    for (size_t k = threadIdx.x; k < bytes / 4; k += blockDim.x) d1[k] = s1[k];
  }
  off += Al16(bytes);
  return dst;
}

template <typename R>
__global__ void __launch_bounds__(kBlock) k_lookup_lds(DevTables<R> T, int ptype, int n, const int32_t* __restrict__ imc,
                                                       const double* __restrict__ ekin, const double* __restrict__ lekin,
                                                       double* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char lds[];
  DevEl<R> e = T.el[ptype];
  const int nmc = T.nmc;
  const size_t N = (size_t)e.N, L = 5 * N * (size_t)nmc, X = (size_t)e.xsN;
  size_t off = 0;
  e.lossGrid64 = static_cast<const double*>(StageArray(lds, off, e.lossGrid64, N * sizeof(double)));
  e.lossData64 = static_cast<const double*>(StageArray(lds, off, e.lossData64, L * sizeof(double)));
  e.xsData64 = static_cast<const double*>(StageArray(lds, off, e.xsData64, X * sizeof(double)));
  e.xsBlock = static_cast<const XsBlock<R>*>(StageArray(lds, off, e.xsBlock, 2 * (size_t)nmc * sizeof(XsBlock<R>)));
  if constexpr (sizeof(R) == sizeof(double)) {
    // the FP64 form's value arrays are the double arrays
    e.lossGrid = reinterpret_cast<const R*>(e.lossGrid64);
    e.lossData = reinterpret_cast<const R*>(e.lossData64);
    e.xsData = reinterpret_cast<const R*>(e.xsData64);
  } else {
    e.lossGrid = static_cast<const R*>(StageArray(lds, off, e.lossGrid, N * sizeof(R)));
    e.lossData = static_cast<const R*>(StageArray(lds, off, e.lossData, L * sizeof(R)));
    e.xsData = static_cast<const R*>(StageArray(lds, off, e.xsData, X * sizeof(R)));
  }
  __syncthreads();
  // each block serves one contiguous chunk of `per` queries per thread; the
  // next query's inputs are loaded before this one is computed, so their HBM
  // latency overlaps the table work (the chunk's last round reloads its own
  // query instead, so the loads are unconditional)
  const int64_t threads = (int64_t)gridDim.x * blockDim.x;
  const int64_t per = ((int64_t)n + threads - 1) / threads;
  const int64_t base = (int64_t)blockIdx.x * blockDim.x * per;
  const int64_t end = min((int64_t)n, base + per * (int64_t)blockDim.x);
  const int64_t stride = blockDim.x;
  int64_t i = base + threadIdx.x;
  if (i >= end) return;
  int mc = imc[i];
  double E = ekin[i], lnE = lekin[i];
  while (true) {
    const int64_t next = i + stride;
    const int64_t ld = next < end ? next : i;
    const int mcN = imc[ld];
    const double EN = ekin[ld], lnEN = lekin[ld];
    LookupStore(e, nmc, n, (int)i, mc, E, lnE, out);
    if (next >= end) break;
    i = next;
    mc = mcN;
    E = EN;
    lnE = lnEN;
  }
}

__global__ void k_reset_state(hepem_tracks t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = t.n;
  if (i >= n) return;
#pragma unroll
  for (int k = 0; k < 4; ++k) t.num_ia_left[k * n + i] = -1.0;
  t.msc_init_range[i] = 1.0e+21;
  t.msc_dyn_range_factor[i] = 0.04;
  t.msc_tlimit_min[i] = 1.0E-7;
  t.msc_first_step[i] = 1;
  t.rng_ctr[i] = 0u;
  t.rng_gauss[i] = 0.0;
  t.rng_has_gauss[i] = 0;
}

}  // namespace hepem
