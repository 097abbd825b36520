// hepem_shower.hip -- showers in a 1D stack of slabs (include/g4hepem_amd_shower.h).
//
// The track stack is a pair of structure-of-arrays buffers of `capacity`
// slots. One stepping round over the n tracks in flight:
//   k_sh_el_substep  the e-/e+ (the first part of the stack): one sub-step of
//                  G4HepEmTrackingManager::TrackElectron's step each, after
//                  the step's pre-step part where a track is not within a step
//                  (device/dev_tracking.hh), with the slab as the geometry and
//                  the MSC displacement applied; a track within its step
//                  carries the loop state to the next round
//   k_sh_el_end    the e-/e+ whose step ended: the loss fluctuation, at-rest
//                  annihilation or the discrete interaction
//   k_sh_gamma     the gammas: HowFar, the cut at the next plane along x (or,
//                  with Woodcock tracking on and above its energy limit, a
//                  step through the stack to its next interaction point or out
//                  of it), then the boundary update or SelectInteraction +
//                  Perform
//   k_sh_post      slab crossing, leakage, scoring, the number of stack slots
//                  each track needs next round and their exclusive scan within
//                  the block
//   k_scan_sums    exclusive scans of the block totals (one block)
//   k_sh_scatter   survivors and their secondaries into the other buffer:
//                  e-/e+ first, then gammas, each in slot order of the parents
//                  (deterministic), planar with the new count
//   k_sh_inject    when the stack runs low, the primaries of the next events
//                  after them.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstring>
#include <limits>
#include <new>
#include <vector>

#include "device/dev_physics.hh"
#include "device/dev_scan.hh"
#include "device/dev_soa.hh"
#include "device/dev_tracking.hh"
#include "g4hepem_amd_shower.h"
#include "hepem_internal.hh"

namespace {

constexpr int kB = 256;
constexpr double kFix = 1.0e9;           // fixed-point units per MeV
constexpr double kGeomMinLength = 5.0e-8;
constexpr double kElectronMass = 0.51099895000;
constexpr int kMaxRounds = 1 << 20;       // stepping rounds without new primaries

#define SH_OK(call)                                   \
  do {                                                \
    if ((call) != hipSuccess) return HEPEM_ERR_HIP;   \
  } while (0)

// splitmix64 finaliser: the stream selector of a primary and of every secondary
__host__ __device__ inline uint64_t Mix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t PrimaryStream(int event) { return Mix64(0x5EED000000000000ull ^ (uint64_t)event); }
__host__ __device__ inline uint64_t ChildStream(uint64_t parent, int nstep, int slot) {
  return Mix64(parent ^ Mix64(((uint64_t)(uint32_t)nstep << 2) | (uint64_t)(slot + 1)));
}

// slab geometry, passed by value
struct Slabs {
  int nbins;          // nlayers * nabs
  int nabs;
  int imc[HEPEM_SLAB_MAX_ABS];
  const double* plane;  // nbins + 1 plane positions [mm], plane[0] = 0
  // per absorber: production thresholds (e-, e+, gamma) of its couple and
  // whether its region applies them to secondaries
  double cutE[HEPEM_SLAB_MAX_ABS], cutP[HEPEM_SLAB_MAX_ABS], cutG[HEPEM_SLAB_MAX_ABS];
  int applyCuts[HEPEM_SLAB_MAX_ABS];
  // Woodcock tracking: gammas above wdtMin (0: off) through the whole stack,
  // whose exit planes are 0 and xmax, with the couple wdtImc of its densest
  // material wdtMat; mat[a] = material of absorber a
  double wdtMin, xmax;
  int wdtImc, wdtMat;
  int mat[HEPEM_SLAB_MAX_ABS];
  unsigned wdtMask;  // bit a: absorber a is of material wdtMat
};

static_assert(HEPEM_SLAB_MAX_ABS == 4, "AbsorberCouple selects over four absorbers");
// g.imc[a] as selects (a per-lane index into the kernel argument would be a
// memory load in the innermost loop)
__device__ inline int AbsorberCouple(const Slabs& g, int a) {
  return a == 0 ? g.imc[0] : (a == 1 ? g.imc[1] : (a == 2 ? g.imc[2] : g.imc[3]));
}

// one stack buffer
// What an e-/e+ carries between the sub-steps of one TrackElectron step
// (G4HepEmTrackingManager.cc:427-606), which run one per stepping round: the
// step's pre-step part sets it and each sub-step updates it (k_sh_el_substep),
// the step's end reads it (k_sh_el_end). Gamma slots do not carry it.
constexpr uint8_t kInStep = 1;   // between sub-steps of a step (else the next round begins a step)
constexpr uint8_t kCont = 2;     // continueStepping: MSC may limit the next sub-step too
constexpr uint8_t kStopped = 4;  // the step's last sub-step stopped the track
constexpr uint8_t kLeft = 8;     // the step's last sub-step left the world
struct ElLoop {
  double* range;       // restricted range, reduced over the sub-steps
  double* stepLeft;    // what is left of the discrete step
  double* mfp;         // [4][n] macroscopic mean free paths of the step's pre-step point
  double* preEkin;     // the step's pre-step energy
  double* meanLoss;    // summed mean energy losses
  int32_t* nsub;       // sub-steps taken in the step
  int8_t* iDProc;      // the discrete process HowFarToDiscreteInteraction chose
  int8_t* winner;      // the process that limited the last sub-step
  uint8_t* flags;      // kInStep | kCont | kStopped | kLeft
};

struct Stack {
  hepem_tracks t;     // t.n = tracks in it (stride of the planar fields)
  double* x;
  int32_t* layer;     // slab index
  int32_t* event;     // deposit row of the event (events resident at once)
  int32_t* nstep;     // steps taken by the track
  ElLoop lp;          // e-/e+ slots: the state between sub-steps
};

__global__ void k_sh_fill64(int64_t* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Primaries of events e0 + j, j in [0, k): at x = 0, moving along +x,
// entering slab 0; deposit row ev0 + j. They go to slots base[0] + j of a
// buffer of total[0] tracks (base = total = NULL: slots j of a buffer of k).
// Writes nothing if the buffer would exceed `cap`.
__global__ void k_sh_inject(Stack s, Slabs g, int ptype, double ekin, int e0, int ev0, int k,
                            const uint32_t* __restrict__ base, const uint32_t* __restrict__ total, int cap) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= k) return;
  const int m = total ? (int)*total : k;
  if (m > cap) return;
  const int i = (base ? (int)*base : 0) + j;
  const hepem_tracks& t = s.t;
  const uint64_t sid = PrimaryStream(e0 + j);
  t.ptype[i] = (int8_t)ptype;
  t.id[i] = (int32_t)(uint32_t)sid;
  t.rng_stream[i] = (uint32_t)(sid >> 32);
  t.imc[i] = g.imc[0];
  t.ekin[i] = ekin;
  t.dir[i] = 1.0;
  t.dir[m + i] = 0.0;
  t.dir[2 * m + i] = 0.0;
  t.safety[i] = 0.0;
  t.on_boundary[i] = 1;
  for (int c = 0; c < 4; ++c) t.num_ia_left[c * m + i] = -1.0;
  t.msc_init_range[i] = 1.0e+21;
  t.msc_dyn_range_factor[i] = 0.04;
  t.msc_tlimit_min[i] = 1.0E-7;
  t.msc_first_step[i] = 1;
  t.rng_ctr[i] = 0u;
  t.rng_gauss[i] = 0.0;
  t.rng_has_gauss[i] = 0;
  s.x[i] = 0.0;
  s.layer[i] = 0;
  s.event[i] = ev0 + j;
  s.nstep[i] = 0;
  s.lp.flags[i] = 0;
}

// isotropic safety of a point in slab l: distance to the nearer plane
__device__ inline double Safety(const Slabs& g, int l, double x) {
  return fmax(0.0, fmin(x - g.plane[l], g.plane[l + 1] - x));
}

// Cuts the proposed geometrical step at the next plane along the pre-step
// direction; xend = end point of the straight move, side = +1/-1 when the
// step ends on the upper/lower plane of the slab, 0 inside.
// Moves (l, a, lo = plane[l], hi = plane[l + 1]) to the slab holding x:
// plane[l] <= x < plane[l + 1], clamped to [0, nbins); a = l % nabs. The
// planes increase, so the walk ends at the one slab a bisection would find;
// a gamma's delta step crosses few slabs, so it reads few planes.
__device__ inline void WalkToSlab(const Slabs& g, double x, int& l, int& a, double& lo, double& hi) {
  while (true) {
    if (l < g.nbins - 1 && x >= hi) {
      ++l;
      a = a + 1 == g.nabs ? 0 : a + 1;
      lo = hi;
      hi = g.plane[l + 1];
    } else if (l > 0 && x < lo) {
      --l;
      a = a == 0 ? g.nabs - 1 : a - 1;
      hi = lo;
      lo = g.plane[l];
    } else {
      break;
    }
  }
}

constexpr int kWdtMaxIter = 1 << 20;  // delta steps of one Woodcock step (then: taken to the exit)

// Woodcock tracking of a gamma through the stack
// (G4HepEmWoodcockHelper::KeepTracking, G4HepEm/G4HepEm/src/G4HepEmWoodcockHelper.cc:145-300,
// and its driver, G4HepEmTrackingManager.cc:955-1062), after its HowFar. The
// distance to the stack's exit less 1e-3 mm is eaten up in steps sampled with
// the densest material's total cross section Sigma_w; a point in another
// material is an interaction with probability Sigma/Sigma_w, one in the densest
// material always. On an interaction the gamma moves there (layer l, position
// xe) with the couple, total mean free path and PE cross section of the point;
// on reaching the exit it ends on the stack's boundary (side sd). Either way
// the number of interaction lengths left is reset and the step adds nothing to
// it. Returns false (nothing changed) if the gamma is within 1e-3 + 1e-6 mm of
// the exit (FindWDTVolume, :141-142).
template <typename R>
__device__ bool WoodcockStep(const hepem::DevTables<R>& T, const Slabs& g, hepem::Trk<R>& gt, hepem::Rng& rng, R& pe, int& l,
                             double x, double ux, double& xe, int& sd) {
  using namespace hepem;
  const double distOut = ux > 0.0 ? (g.xmax - x) / ux : (ux < 0.0 ? (0.0 - x) / ux : 1.0e300);
  double dist = distOut - 1.0E-3 > 0.0 ? distOut - 1.0E-3 : 0.0;
  if (!(dist >= 1.0E-6)) return false;
  constexpr R kBig = std::numeric_limits<R>::max();
  Trk<R> xt;  // cross-section evaluations at the gamma's energy
  xt.isEl = false;
  xt.setE(gt.ekin);
  xt.imc = g.wdtImc;
  R wpe = (R)0.0;
  const R wmx = GetTotalMacXSec(T, xt, wpe);
  const R wmfp = wmx > (R)0.0 ? (R)1.0 / wmx : kBig;
  // total and PE cross sections of the other absorbers' couples at this energy
  // (the energy stays the same through the step), ahead of the loop so that
  // its lanes do not diverge over them
  R mxa[HEPEM_SLAB_MAX_ABS], pea[HEPEM_SLAB_MAX_ABS];
#pragma unroll
  for (int k = 0; k < HEPEM_SLAB_MAX_ABS; ++k) {
    mxa[k] = pea[k] = (R)0.0;
    if (k < g.nabs && !((g.wdtMask >> k) & 1u)) {
      xt.imc = g.imc[k];
      mxa[k] = GetTotalMacXSec(T, xt, pea[k]);
    }
  }
  R mfp = wmfp, pint = wpe;
  int lp = l, ap = l % g.nabs, imcp = g.wdtImc;
  double plo = g.plane[l], phi = g.plane[l + 1];
  bool boundary = false;
  double L = 0.0;
  for (int it = 0;; ++it) {
    if (it >= kWdtMaxIter) {
      L += dist;
      boundary = true;
      break;
    }
    const R pstep = wmfp < kBig ? -vlog((R)rng.flat()) * wmfp : kBig;
    if (dist < (double)pstep) {
      L += dist;
      boundary = true;
      break;
    }
    L += (double)pstep;
    dist -= (double)pstep;
    WalkToSlab(g, x + L * ux, lp, ap, plo, phi);
    if (!((g.wdtMask >> ap) & 1u)) {
      const R mx = ap == 0 ? mxa[0] : (ap == 1 ? mxa[1] : (ap == 2 ? mxa[2] : mxa[3]));
      if ((double)(mx * wmfp) > rng.flat()) {
        imcp = AbsorberCouple(g, ap);
        mfp = mx > (R)0.0 ? (R)1.0 / mx : kBig;
        pint = ap == 0 ? pea[0] : (ap == 1 ? pea[1] : (ap == 2 ? pea[2] : pea[3]));
        break;
      }
    } else {
      imcp = g.wdtImc;
      mfp = wmfp;
      pint = wpe;
      break;
    }
  }
  gt.nial[0] = (R)-1.0;
  gt.gstep = (R)0.0;
  if (boundary) {
    sd = ux < 0.0 ? -1 : 1;
    l = sd > 0 ? g.nbins - 1 : 0;
    gt.onBoundary = true;
    xe = sd > 0 ? g.xmax : 0.0;
  } else {
    l = lp;
    gt.imc = imcp;
    gt.mfp[0] = mfp;
    pe = pint;
    gt.onBoundary = false;
    sd = 0;
    xe = x + L * ux;
  }
  return true;
}

// The e-/e+ slots [0, ne) of G4HepEmTrackingManager::TrackElectron's step with
// the slab as the caller's geometry (device/dev_tracking.hh), one sub-step
// per stepping round: every slot takes one sub-step, after beginning a step
// if it is not within one (k_sh_el_substep), and a slot whose step thereby
// ended finishes it (k_sh_el_end). A slot within a step stays in the stack
// with its loop state, deposits nothing and creates nothing until its step
// ends. Every round thus does one uniform unit of work per e-/e+, whatever the
// length of its sub-step chain.
//
// k_sh_el_substep: for a slot not within a step, the pre-step safety (kept in
// t.safety for the step), the interaction lengths and
// HowFarToDiscreteInteraction; then one sub-step (SlabSubstep: HowFarToMSC,
// the move, the along-step part, the MSC displacement). xend = post-sub-step
// position; side = +1/-1 when it ended on the upper or lower plane of its
// slab, 0 inside. A slot still within its step gets no deposit or
// secondaries this round.
template <typename R>
__global__ void __launch_bounds__(kB) k_sh_el_substep(Stack s, Slabs g, hepem_step_out res, double* __restrict__ xend,
                                                      int8_t* __restrict__ side, hepem::DevTables<R> T, uint64_t seed,
                                                      int ne) {
  using namespace hepem;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ne) return;
  const int n = s.t.n;
  hepem_tracks& t = s.t;
  const int l = s.layer[i];
  ElState<R> el;
  Rng rng;
  InitElectron(t, n, i, seed, el, rng);
  StepLoop<R> L;
  SlabPos p{s.x[i], g.plane[l], g.plane[l + 1], l == 0, l == g.nbins - 1, 0};
  const uint8_t f0 = s.lp.flags[i];
  if (!(f0 & kInStep)) {
    SlabBegin(T, el, L, rng, p);
    t.safety[i] = (double)el.t.safety;
    s.lp.preEkin[i] = L.preEkin;
    s.lp.iDProc[i] = (int8_t)L.iDProc;
#pragma unroll
    for (int k = 0; k < 4; ++k) s.lp.mfp[k * n + i] = (double)el.t.mfp[k];
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) el.t.mfp[k] = (R)s.lp.mfp[k * n + i];
    el.range = (R)s.lp.range[i];
    el.pstep = (R)s.lp.stepLeft[i];
    el.t.gstep = el.pstep;  // HowFarToMSC keeps it where MSC is off
    el.preEkin = el.t.ekin;
    el.preLEkin = el.t.logE();
    L.iDProc = s.lp.iDProc[i];
    el.t.winner = L.iDProc;
    L.stepLeft = el.pstep;
    L.truePath = (R)0.;
    L.meanLoss = s.lp.meanLoss[i];
    L.preEkin = s.lp.preEkin[i];
    L.preLEkin = (R)0.;
    L.cont = (f0 & kCont) != 0;
    L.stopped = false;
    L.left = false;
    L.nsub = s.lp.nsub[i];
  }
  const bool done = SlabSubstep(T, el, L, rng, p);
  StoreElectronState(t, n, i, el, rng);
  t.ekin[i] = el.t.ekin;
#pragma unroll
  for (int c = 0; c < 3; ++c) t.dir[c * n + i] = (double)el.t.dir[c];
  t.on_boundary[i] = p.side != 0 ? 1 : 0;
  s.lp.range[i] = (double)el.range;
  s.lp.stepLeft[i] = (double)L.stepLeft;
  s.lp.meanLoss[i] = L.meanLoss;
  s.lp.nsub[i] = L.nsub;
  s.lp.winner[i] = (int8_t)el.t.winner;
  s.lp.flags[i] = done ? (uint8_t)((L.stopped ? kStopped : 0) | (L.left ? kLeft : 0)) : (uint8_t)(kInStep | kCont);
  if (!done) {
    res.edep[i] = 0.0;
    res.nsec[i] = 0;
  }
  xend[i] = p.x;
  side[i] = (int8_t)p.side;
}

// k_sh_el_end: the step's end for the slots whose step ended this round: the
// loss fluctuation on the summed mean loss, at-rest annihilation, or the choice
// of the discrete interaction and its final state (LoopFinish).
template <typename R>
__global__ void __launch_bounds__(kB) k_sh_el_end(Stack s, hepem_step_out res, hepem::DevTables<R> T, uint64_t seed, int ne) {
  using namespace hepem;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ne) return;
  const uint8_t f = s.lp.flags[i];
  if (f & kInStep) return;
  const int n = s.t.n;
  hepem_tracks& t = s.t;
  ElState<R> el;
  Rng rng;
  InitElectron(t, n, i, seed, el, rng);
#pragma unroll
  for (int k = 0; k < 4; ++k) el.t.mfp[k] = (R)s.lp.mfp[k * n + i];
  el.range = (R)s.lp.range[i];
  el.t.winner = s.lp.winner[i];
  StepLoop<R> L;
  L.iDProc = s.lp.iDProc[i];
  L.meanLoss = s.lp.meanLoss[i];
  L.preEkin = s.lp.preEkin[i];
  L.preLEkin = L.preEkin > 0. ? vlog((R)L.preEkin) : (R)-30;
  L.stopped = (f & kStopped) != 0;
  L.left = (f & kLeft) != 0;
  L.cont = false;
  SecList<R> sl;
  sl.ns = 0;
  const int p = LoopFinishSelect(T, el, L, sl, rng);
  res.edep[i] = el.t.edep;
  res.winner[i] = el.t.winner;
  ElectronFinalState(T, el.t, p, sl, rng);
  t.ekin[i] = el.t.ekin;
#pragma unroll
  for (int c = 0; c < 3; ++c) t.dir[c * n + i] = (double)el.t.dir[c];
#pragma unroll
  for (int k = 0; k < 4; ++k) t.num_ia_left[k * n + i] = (double)el.t.nial[k];
  StoreRng(t, i, rng);
  WriteSecondaries(res, n, i, sl);
}

// One step of the gamma slots [ne, n), as G4HepEmTrackingManager::TrackGamma
// takes it (G4HepEmTrackingManager.cc:893-1145): HowFar, the cut at the next
// plane along x (or, with Woodcock tracking on and above its energy limit, a
// Woodcock step through the stack), then the boundary update or
// SelectInteraction + Perform. xend and side as k_sh_el_substep; layer = the slab
// the step's deposit is scored in (changed only by a Woodcock step).
template <typename R, bool kWoodcock>
__global__ void __launch_bounds__(kB) k_sh_gamma(Stack s, Slabs g, hepem_step_out res, double* __restrict__ xend,
                                                 int8_t* __restrict__ side, hepem::DevTables<R> T, uint64_t seed, int ne) {
  using namespace hepem;
  const int i = ne + blockIdx.x * blockDim.x + threadIdx.x;
  const int n = s.t.n;
  if (i >= n) return;
  hepem_tracks& t = s.t;
  int l = s.layer[i];
  const double x = s.x[i];
  SecList<R> sl;
  sl.ns = 0;
  Trk<R> gt;
  gt.isEl = false;
  LoadTrack(t, n, i, gt);
  Rng rng;
  rng.init(seed, (uint32_t)t.id[i], StreamHi(t, i), t.rng_ctr[i], t.rng_has_gauss[i] != 0, t.rng_gauss[i]);
  R pe = GammaHowFar(T, gt, rng);
  const double ux = (double)gt.dir[0];
  double xe;
  int sd = 0;
  if (!(kWoodcock && gt.ekin > g.wdtMin && WoodcockStep(T, g, gt, rng, pe, l, x, ux, xe, sd))) {
    const double gs = (double)gt.gstep;
    double dist = 1.0e300;
    if (ux > 0.0) {
      dist = (g.plane[l + 1] - x) / ux;
      sd = 1;
    } else if (ux < 0.0) {
      dist = (g.plane[l] - x) / ux;
      sd = -1;
    }
    if (dist < gs) {
      gt.gstep = (R)dist;
      gt.onBoundary = true;
      xe = sd > 0 ? g.plane[l + 1] : g.plane[l];
    } else {
      gt.onBoundary = false;
      sd = 0;
      xe = fmin(fmax(x + gs * ux, g.plane[l]), g.plane[l + 1]);
    }
  } else {
    s.layer[i] = l;
    t.imc[i] = gt.imc;
  }
  GammaPerform(T, gt, pe, sl, rng);
  t.ekin[i] = gt.ekin;
#pragma unroll
  for (int c = 0; c < 3; ++c) t.dir[c * n + i] = (double)gt.dir[c];
  t.num_ia_left[i] = (double)gt.nial[0];
  StoreRng(t, i, rng);
  t.on_boundary[i] = gt.onBoundary ? 1 : 0;
  res.true_step[i] = (double)gt.gstep;
  res.edep[i] = gt.edep;
  res.winner[i] = gt.winner;
  WriteSecondaries(res, n, i, sl);
  xend[i] = xe;
  side[i] = (int8_t)sd;
}

// After Perform: position, slab crossing, scoring and next-round slot counts.
// The next stack holds the e-/e+ first and the gammas after them (so the
// library's particle-type queue covers runs of consecutive slots), each part
// in slot order of the parents. keep[i] = 1 if the track itself goes on,
// bit s of smask[i] = secondary slot s is stacked; offE/offG[i] = exclusive scans within the
// block of the e-/e+ and gamma slots track i needs; bsumE/bsumG/bsec[block] =
// the block's totals of those and of its secondaries. Deposits are summed per
// block and (event, slab) before the global atomic; the fixed-point sum does
// not depend on the grouping.
__global__ void __launch_bounds__(kB) k_sh_post(Stack s, Slabs g, hepem_step_out res, const double* __restrict__ xend,
                          const int8_t* __restrict__ side, unsigned long long* dep, unsigned long long* leak,
                          uint8_t* keep, uint8_t* smask, uint32_t* offE, uint32_t* offG, uint32_t* bsumE, uint32_t* bsumG,
                          uint32_t* bsec) {
  const int i = blockIdx.x * kB + threadIdx.x;
  const int n = s.t.n;
  const bool valid = i < n;
  const hepem_tracks& t = s.t;
  // the steps that ended this round (an e-/e+ within its sub-steps has not),
  // summed with the block's stacked secondaries in bsec (bits 16 and up)
  const uint32_t ended = (valid && (t.ptype[i] == 2 || !(s.lp.flags[i] & kInStep))) ? 1u : 0u;
  constexpr unsigned long long kNone = ~0ull;
  unsigned long long key = kNone, val = 0ull;
  uint32_t cE = 0u, cG = 0u, ns = 0u;
  if (valid) {
    const int ev = s.event[i];
    const int l0 = s.layer[i];  // the pre-step slab: scoring and the couple's cuts
    int l = l0;
    double edep = res.edep[i];
    const double x = xend[i];
    const int sd = side[i];
    const bool isE = t.ptype[i] != 2;
    const double ekin = t.ekin[i];
    bool alive = ekin > 0.0;
    int nsec = res.nsec[i];
    if (sd != 0 && (l + sd < 0 || l + sd >= g.nbins)) {
      // leaves the stack: its kinetic energy, the annihilation energy of an e+
      // in flight and anything its last step created are lost
      double e = ekin + (t.ptype[i] == 1 && ekin > 0.0 ? 2.0 * kElectronMass : 0.0);
      for (int sl = 0; sl < nsec; ++sl) e += res.sec_ekin[sl * n + i];
      atomicAdd(&leak[ev], (unsigned long long)llrint(e * kFix));
      alive = false;
      nsec = 0;
    } else if (alive && sd != 0) {
      l += sd;
    }
    s.x[i] = x;
    s.layer[i] = l;
    // StackSecondaries (G4HepEmTrackingManager.cc:1274-1313): where the
    // couple's region applies cuts, a secondary below its production
    // threshold is deposited at the step instead of stacked (an e+ with its
    // annihilation energy, when the gamma threshold exceeds the electron mass)
    const int a = l0 % g.nabs;
    uint32_t mask = 0u;
    for (int sl = 0; sl < nsec; ++sl) {
      const int8_t pt = res.sec_ptype[sl * n + i];
      const double se = res.sec_ekin[sl * n + i];
      bool cut = false;
      if (g.applyCuts[a]) {
        if (pt == 0) cut = se < g.cutE[a];
        else if (pt == 1) cut = kElectronMass < g.cutG[a] && se < g.cutP[a];
        else cut = se < g.cutG[a];
      }
      if (cut) {
        edep += pt == 1 ? se + 2.0 * kElectronMass : se;
      } else {
        mask |= 1u << sl;
        ++ns;
        if (pt == 2) ++cG;
        else ++cE;
      }
    }
    keep[i] = alive ? 1 : 0;
    smask[i] = (uint8_t)mask;
    if (alive) {
      if (isE) ++cE;
      else ++cG;
    }
    if (edep > 0.0) {
      key = (unsigned long long)ev * (unsigned long long)g.nbins + (unsigned long long)l0;
      val = (unsigned long long)llrint(edep * kFix);
    }
  }
  // per-block aggregation of the deposits in an LDS hash table (open
  // addressing, at most kB keys in 2 kB entries, so every probe sequence ends):
  // one global atomic per distinct (event, slab) of the block
  constexpr int kHash = 2 * kB;
  __shared__ unsigned long long hkey[kHash], hval[kHash];
  for (int e = threadIdx.x; e < kHash; e += kB) {
    hkey[e] = kNone;
    hval[e] = 0ull;
  }
  __syncthreads();
  if (key != kNone) {
    uint32_t h = (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> 55) & (kHash - 1);
    while (true) {
      const unsigned long long prev = atomicCAS(&hkey[h], kNone, key);
      if (prev == kNone || prev == key) {
        atomicAdd(&hval[h], val);
        break;
      }
      h = (h + 1) & (kHash - 1);
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kHash; e += kB) {
    const unsigned long long k = hkey[e];
    if (k != kNone) atomicAdd(&dep[k], hval[e]);
  }
  uint32_t xE, xG, tE, tG, tns;
  hepem::BlockScan2<kB>(cE, cG, ns | (ended << 16), xE, xG, tE, tG, tns);
  if (valid) {
    offE[i] = xE;
    offG[i] = xG;
  }
  if (threadIdx.x == 0) {
    bsumE[blockIdx.x] = tE;
    bsumG[blockIdx.x] = tG;
    bsec[blockIdx.x] = tns;
  }
}

// In-place exclusive scans of the nb block sums of e-/e+ and gamma slots by
// one block; the round's secondaries are added to *nsecTotal. The `inject`
// primaries (e-/e+ if injectE) go to the end of their part of the next stack:
// total[0] = tracks of the next stack, total[1] = first slot of the injected
// primaries, total[2] = slots of the e-/e+ part. The track count is also
// written to host memory, followed by `tag` once it is visible there.
constexpr int kScanB = 1024;  // threads of the one-block scan
__global__ void __launch_bounds__(kScanB) k_scan_sums(uint32_t* bsumE, uint32_t* bsumG, const uint32_t* __restrict__ bsec, int nb,
                                                  int inject, int injectE, uint32_t* total, unsigned long long* nsecTotal,
                                                  unsigned long long* nstepsTotal, volatile uint32_t* hostCount, uint32_t tag) {
  __shared__ uint32_t carryE, carryG;
  __shared__ unsigned long long secs, steps;
  if (threadIdx.x == 0) {
    carryE = 0u;
    carryG = 0u;
    secs = 0ull;
    steps = 0ull;
  }
  __syncthreads();
  unsigned long long myEnded = 0ull;
  for (int base = 0; base < nb; base += kScanB) {
    const int i = base + threadIdx.x;
    const uint32_t a = i < nb ? bsumE[i] : 0u;
    const uint32_t b = i < nb ? bsumG[i] : 0u;
    const uint32_t packed = i < nb ? bsec[i] : 0u;  // secondaries | steps ended << 16
    const uint32_t w = packed & 0xFFFFu;
    myEnded += packed >> 16;
    uint32_t xa, xb, ta, tb, tw;
    hepem::BlockScan2<kScanB>(a, b, w, xa, xb, ta, tb, tw);
    if (i < nb) {
      bsumE[i] = carryE + xa;
      bsumG[i] = carryG + xb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      carryE += ta;
      carryG += tb;
      secs += tw;
    }
    __syncthreads();
  }
  for (int o = 32; o > 0; o >>= 1) myEnded += __shfl_xor(myEnded, o, 64);
  if (__lane_id() == 0) atomicAdd(&steps, myEnded);
  __syncthreads();
  if (threadIdx.x == 0) {
    *nstepsTotal += steps;
    const uint32_t k = (uint32_t)inject;
    const uint32_t elen = carryE + (injectE ? k : 0u);
    total[0] = carryE + carryG + k;
    total[1] = injectE ? carryE : elen + carryG;
    total[2] = elen;
    *nsecTotal += secs;
    // the host reads the next round's track count while the scatter runs
    hostCount[0] = carryE + carryG + k;
    hostCount[2] = elen;  // its e-/e+ part
    __threadfence_system();
    hostCount[1] = tag;
  }
}

// Survivor i and its stacked secondaries go to the next e-/e+ slots from
// bsumE[block] + offE[i] and the next gamma slots from total[2] +
// bsumG[block] + offG[i], in the order survivor, secondary 0, secondary 1;
// each starts with the isotropic safety of its position. Writes nothing beyond
// `cap` (the host then reports the overflow).
__global__ void __launch_bounds__(kB) k_sh_scatter(Stack a, Stack b, Slabs g, hepem_step_out res, const uint8_t* __restrict__ keep,
                             const uint8_t* __restrict__ smask, const uint32_t* __restrict__ offE,
                             const uint32_t* __restrict__ offG, const uint32_t* __restrict__ bsumE,
                             const uint32_t* __restrict__ bsumG, const uint32_t* __restrict__ total, int cap) {
  const int i = blockIdx.x * kB + threadIdx.x;
  const int n = a.t.n;
  if (i >= n) return;
  const int m = (int)total[0];
  if (m > cap) return;
  const hepem_tracks& s = a.t;
  const hepem_tracks& d = b.t;
  int jE = (int)(bsumE[blockIdx.x] + offE[i]);
  int jG = (int)(total[2] + bsumG[blockIdx.x] + offG[i]);
  const double x = a.x[i];
  const int l = a.layer[i];
  const double safety = Safety(g, l, x);
  const uint64_t sid = ((uint64_t)s.rng_stream[i] << 32) | (uint32_t)s.id[i];
  const int ns1 = a.nstep[i] + 1;
  // A gamma reads only the first interaction-length slot (its total) and
  // neither the safety nor the MSC state: those fields of a gamma slot are not
  // carried (the e-/e+ and gamma parts of the stack keep whole waves on one
  // side of these branches).
  if (keep[i]) {
    const int8_t pt = s.ptype[i];
    const bool isE = pt != 2;
    const int j = isE ? jE++ : jG++;
    d.ptype[j] = pt;
    d.id[j] = s.id[i];
    d.rng_stream[j] = s.rng_stream[i];
    d.imc[j] = g.imc[l % g.nabs];
    d.ekin[j] = s.ekin[i];
    for (int c = 0; c < 3; ++c) d.dir[c * m + j] = s.dir[c * n + i];
    d.on_boundary[j] = s.on_boundary[i];
    d.num_ia_left[j] = s.num_ia_left[i];
    bool inStep = false;
    if (isE) {
      for (int k = 1; k < 4; ++k) d.num_ia_left[k * m + j] = s.num_ia_left[k * n + i];
      d.msc_init_range[j] = s.msc_init_range[i];
      d.msc_dyn_range_factor[j] = s.msc_dyn_range_factor[i];
      d.msc_tlimit_min[j] = s.msc_tlimit_min[i];
      d.msc_first_step[j] = s.msc_first_step[i];
      const uint8_t f = a.lp.flags[i];
      inStep = (f & kInStep) != 0;
      b.lp.flags[j] = inStep ? f : (uint8_t)0;
      if (inStep) {
        // within a step: its pre-step safety and the loop state go along
        d.safety[j] = s.safety[i];
        b.lp.range[j] = a.lp.range[i];
        b.lp.stepLeft[j] = a.lp.stepLeft[i];
        for (int k = 0; k < 4; ++k) b.lp.mfp[k * m + j] = a.lp.mfp[k * n + i];
        b.lp.preEkin[j] = a.lp.preEkin[i];
        b.lp.meanLoss[j] = a.lp.meanLoss[i];
        b.lp.nsub[j] = a.lp.nsub[i];
        b.lp.iDProc[j] = a.lp.iDProc[i];
      } else {
        d.safety[j] = safety;
      }
    }
    d.rng_ctr[j] = s.rng_ctr[i];
    d.rng_gauss[j] = s.rng_gauss[i];
    d.rng_has_gauss[j] = s.rng_has_gauss[i];
    b.x[j] = x;
    b.layer[j] = l;
    b.event[j] = a.event[i];
    b.nstep[j] = inStep ? a.nstep[i] : ns1;  // a step counts once it ends
  }
  const uint32_t mask = smask[i];
  for (int sl = 0; sl < 2; ++sl) {
    if (!((mask >> sl) & 1u)) continue;
    const uint64_t cid = ChildStream(sid, ns1, sl);
    const int8_t pt = res.sec_ptype[sl * n + i];
    const bool isE = pt != 2;
    const int j = isE ? jE++ : jG++;
    d.ptype[j] = pt;
    d.id[j] = (int32_t)(uint32_t)cid;
    d.rng_stream[j] = (uint32_t)(cid >> 32);
    d.imc[j] = g.imc[l % g.nabs];
    d.ekin[j] = res.sec_ekin[sl * n + i];
    for (int c = 0; c < 3; ++c) d.dir[c * m + j] = res.sec_dir[(c * 2 + sl) * n + i];
    d.on_boundary[j] = 0;
    d.num_ia_left[j] = -1.0;
    if (isE) {
      d.safety[j] = safety;
      b.lp.flags[j] = 0;
      for (int k = 1; k < 4; ++k) d.num_ia_left[k * m + j] = -1.0;
      d.msc_init_range[j] = 1.0e+21;
      d.msc_dyn_range_factor[j] = 0.04;
      d.msc_tlimit_min[j] = 1.0E-7;
      d.msc_first_step[j] = 1;
    }
    d.rng_ctr[j] = 0u;
    d.rng_gauss[j] = 0.0;
    d.rng_has_gauss[j] = 0;
    b.x[j] = x;
    b.layer[j] = l;
    b.event[j] = a.event[i];
    b.nstep[j] = 0;
  }
}

// device memory of one run, freed on every exit path
struct Arena {
  std::vector<void*> ptrs;
  ~Arena() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename T> bool get(T*& p, size_t n) {
    void* v = nullptr;
    if (hipMalloc(&v, (n ? n : 1) * sizeof(T)) != hipSuccess) return false;
    ptrs.push_back(v);
    p = static_cast<T*>(v);
    return true;
  }
};

// Three words of pinned, device-mapped, coherent host memory: the next round's
// track count, the tag of the round that wrote it and the count's e-/e+ part. Freed on every exit
// path (after the stream has drained).
struct HostCount {
  uint32_t* p = nullptr;  // host address
  uint32_t* d = nullptr;  // device address of the same words
  bool alloc() {
    if (hipHostMalloc(reinterpret_cast<void**>(&p), 64, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) return false;
    p[0] = 0u;
    p[1] = 0u;
    p[2] = 0u;
    return hipHostGetDevicePointer(reinterpret_cast<void**>(&d), p, 0) == hipSuccess;
  }
  ~HostCount() {
    if (p) (void)hipHostFree(p);
  }
  // Waits until a kernel has published `tag`; false on a stream error, or if
  // the stream drained without it.
  bool wait(uint32_t tag, hipStream_t st) const {
    for (uint64_t spin = 1;; ++spin) {
      if (__atomic_load_n(&p[1], __ATOMIC_ACQUIRE) == tag) return true;
      if ((spin & 1023u) == 0u) {
        const hipError_t e = hipStreamQuery(st);
        if (e == hipSuccess) return __atomic_load_n(&p[1], __ATOMIC_ACQUIRE) == tag;
        if (e != hipErrorNotReady) return false;
      }
    }
  }
  uint32_t count() const { return __atomic_load_n(&p[0], __ATOMIC_ACQUIRE); }
  uint32_t electrons() const { return __atomic_load_n(&p[2], __ATOMIC_ACQUIRE); }
};

// Synchronises the stream when the run ends, on every path, before the
// buffers its kernels use are freed.
struct StreamDrain {
  hipStream_t s;
  ~StreamDrain() { (void)hipStreamSynchronize(s); }
};

bool AllocStack(Arena& A, Stack& s, int cap) {
  hepem_tracks& t = s.t;
  t.n = 0;
  return A.get(t.ptype, cap) && A.get(t.id, cap) && A.get(t.imc, cap) && A.get(t.ekin, cap) && A.get(t.dir, 3 * (size_t)cap) &&
         A.get(t.safety, cap) && A.get(t.on_boundary, cap) && A.get(t.num_ia_left, 4 * (size_t)cap) &&
         A.get(t.msc_init_range, cap) && A.get(t.msc_dyn_range_factor, cap) && A.get(t.msc_tlimit_min, cap) &&
         A.get(t.msc_first_step, cap) && A.get(t.rng_ctr, cap) && A.get(t.rng_gauss, cap) && A.get(t.rng_has_gauss, cap) &&
         A.get(t.rng_stream, cap) && A.get(s.x, cap) && A.get(s.layer, cap) && A.get(s.event, cap) && A.get(s.nstep, cap) &&
         A.get(s.lp.range, cap) && A.get(s.lp.stepLeft, cap) && A.get(s.lp.mfp, 4 * (size_t)cap) && A.get(s.lp.preEkin, cap) &&
         A.get(s.lp.meanLoss, cap) && A.get(s.lp.nsub, cap) && A.get(s.lp.iDProc, cap) && A.get(s.lp.winner, cap) &&
         A.get(s.lp.flags, cap);
}

// Kahan-compensated accumulator
struct Kahan {
  double s = 0.0, c = 0.0;
  void add(double v) {
    const double y = v - c;
    const double t = s + y;
    c = (t - s) - y;
    s = t;
  }
};

}  // namespace

extern "C" int hepem_slab_run(const hepem_ctx* ctx, const hepem_slab_setup* su, int32_t nevents, uint64_t seed,
                              hepem_slab_result* r, void* stream) {
  if (!ctx || !su || !r || !r->edep_sum || !r->edep_sum2 || nevents < 0) return HEPEM_ERR_INVALID;
  if (su->nlayers < 1 || su->nabs < 1 || su->nabs > HEPEM_SLAB_MAX_ABS || su->capacity < 1 || su->events_per_chunk < 1 ||
      !(su->ekin0 > 0.0) || su->primary < 0 || su->primary > 2 || su->refill_below < 0 || !(su->woodcock_min_ekin >= 0.0))
    return HEPEM_ERR_INVALID;
  for (int a = 0; a < su->nabs; ++a)
    if (!(su->thickness[a] > 0.0) || su->imc[a] < 0) return HEPEM_ERR_INVALID;
  const int nbins = su->nlayers * su->nabs;
  const int cap = su->capacity;
  const int chunk = su->events_per_chunk < cap ? su->events_per_chunk : cap;
  const hipStream_t st = static_cast<hipStream_t>(stream);

  std::vector<double> planes((size_t)nbins + 1);
  planes[0] = 0.0;
  for (int k = 0; k < nbins; ++k) planes[k + 1] = planes[k] + su->thickness[k % su->nabs];

  // per-event deposit rows resident at once (events beyond are run afterwards)
  constexpr int64_t kDepBytes = int64_t(1) << 30;
  const int64_t rowBytes = (int64_t)nbins * (int64_t)sizeof(unsigned long long);
  int64_t span = kDepBytes / rowBytes;
  if (span < chunk) span = chunk;
  if (span > nevents) span = nevents;
  if (span < 1) span = 1;
  const int refill = su->refill_below;
  int growRounds = -1;  // rounds from stacking the first events to their largest stack (-1: not yet seen)

  Arena A;
  Stack s[2];
  hepem_step_out res;  // the fields the stepping kernel writes and the post kernel reads
  res.geom_step = nullptr;
  res.displacement = nullptr;
  res.stopped = nullptr;
  double *dPlanes, *xend;
  int8_t* side;
  uint8_t *keep, *smask;
  uint32_t *offs, *bsum, *bsec, *total;  // offs: e-/e+ then gamma block offsets (2 x cap); bsum: 2 x nblk
  unsigned long long *dep, *leak, *nsec;
  const int nblk = (cap + kB - 1) / kB;
  if (!AllocStack(A, s[0], cap) || !AllocStack(A, s[1], cap) || !A.get(res.true_step, cap) || !A.get(res.edep, cap) ||
      !A.get(res.winner, cap) || !A.get(res.nsec, cap) || !A.get(res.sec_ptype, 2 * (size_t)cap) ||
      !A.get(res.sec_ekin, 2 * (size_t)cap) || !A.get(res.sec_dir, 6 * (size_t)cap) || !A.get(dPlanes, planes.size()) ||
      !A.get(xend, cap) || !A.get(side, cap) || !A.get(keep, cap) || !A.get(smask, cap) || !A.get(offs, 2 * (size_t)cap) ||
      !A.get(bsum, 2 * (size_t)nblk) || !A.get(bsec, nblk) || !A.get(total, 3) || !A.get(dep, (size_t)span * nbins) ||
      !A.get(leak, (size_t)span) || !A.get(nsec, 2))
    return HEPEM_ERR_ALLOC;
  HostCount hc;  // next round's track count, published by k_scan_sums
  if (!hc.alloc()) return HEPEM_ERR_ALLOC;
  StreamDrain drain{st};  // declared last: drains the stream before hc and A are freed
  uint32_t tag = 0;
  SH_OK(hipMemcpy(dPlanes, planes.data(), planes.size() * sizeof(double), hipMemcpyHostToDevice));

  Slabs g;
  g.nbins = nbins;
  g.nabs = su->nabs;
  for (int a = 0; a < HEPEM_SLAB_MAX_ABS; ++a) {
    g.imc[a] = a < su->nabs ? su->imc[a] : 0;
    g.cutE[a] = g.cutP[a] = g.cutG[a] = 0.0;
    g.applyCuts[a] = 0;
    if (a < su->nabs) {
      double cuts[3];
      if (HepemCoupleCuts(ctx, su->imc[a], cuts, &g.applyCuts[a]) != HEPEM_OK) return HEPEM_ERR_INVALID;
      g.cutE[a] = cuts[0];
      g.cutP[a] = cuts[1];
      g.cutG[a] = cuts[2];
    }
  }
  g.plane = dPlanes;
  // Woodcock tracking: the stack is one envelope whose densest material (the
  // first absorber of the largest density, G4HepEmWoodcockHelper::FindWDTMaterial,
  // G4HepEmWoodcockHelper.cc:318-338) sets the majorant cross section
  g.wdtMin = su->woodcock_min_ekin;
  g.xmax = planes[(size_t)nbins];
  g.wdtImc = su->imc[0];
  g.wdtMat = -1;
  {
    double maxDensity = -1.0;
    for (int a = 0; a < HEPEM_SLAB_MAX_ABS; ++a) {
      g.mat[a] = -1;
      if (a >= su->nabs) continue;
      double density;
      if (HepemCoupleMaterial(ctx, su->imc[a], &g.mat[a], &density) != HEPEM_OK) return HEPEM_ERR_INVALID;
      if (density > maxDensity) {
        maxDensity = density;
        g.wdtMat = g.mat[a];
        g.wdtImc = su->imc[a];
      }
    }
    g.wdtMask = 0u;
    for (int a = 0; a < su->nabs; ++a)
      if (g.mat[a] == g.wdtMat) g.wdtMask |= 1u << a;
  }
  const bool woodcock = g.wdtMin > 0.0;
  const bool fp64 = HepemPrecision(ctx) == HEPEM_FP64;

  std::vector<Kahan> ks(nbins), ks2(nbins);
  Kahan kleak, ktot;
  std::vector<unsigned long long> hdep((size_t)span * nbins), hleak((size_t)span);
  r->steps = 0;
  r->tracks = 0;
  r->substeps = 0;
  r->iterations = 0;
  r->max_in_flight = 0;
  unsigned long long hnsec[2] = {0ull, 0ull};  // secondaries stacked, steps ended
  hipLaunchKernelGGL(k_sh_fill64, dim3(1), dim3(kB), 0, st, reinterpret_cast<int64_t*>(nsec), 2);

  for (int64_t s0 = 0; s0 < nevents; s0 += span) {
    const int ns = (int)((nevents - s0) < span ? (nevents - s0) : span);
    hipLaunchKernelGGL(k_sh_fill64, dim3((unsigned)(((int64_t)ns * nbins + kB - 1) / kB)), dim3(kB), 0, st,
                       reinterpret_cast<int64_t*>(dep), ns * nbins);
    hipLaunchKernelGGL(k_sh_fill64, dim3((ns + kB - 1) / kB), dim3(kB), 0, st, reinterpret_cast<int64_t*>(leak), ns);
    int next = 0;  // next event of this span to stack
    int n = 0;     // tracks in flight
    int ne = 0;    // the e-/e+ among them (the first ne slots)
    int cur = 0;
    int rounds = 0;
    int sinceInject = 0;  // rounds stepped since primaries were last stacked
    while (true) {
      if (n == 0) {
        if (next >= ns) break;
        const int k = (ns - next) < chunk ? (ns - next) : chunk;
        s[cur].t.n = k;
        hipLaunchKernelGGL(k_sh_inject, dim3((k + kB - 1) / kB), dim3(kB), 0, st, s[cur], g, su->primary, su->ekin0,
                           (int)(s0 + next), next, k, nullptr, nullptr, cap);
        SH_OK(hipGetLastError());
        r->tracks += k;
        next += k;
        n = k;
        ne = su->primary != 2 ? k : 0;
        rounds = 0;
        sinceInject = 0;
      }
      // every step of an e-/e+ loses energy and every gamma step ends on a plane
      // or in an interaction, so the tracks in flight run out; the bound only
      // turns a broken invariant into an error instead of an endless loop
      if (++rounds > kMaxRounds) return HEPEM_ERR_CAPACITY;
      if (n > r->max_in_flight) r->max_in_flight = n;
      Stack& a = s[cur];
      Stack& b = s[cur ^ 1];
      a.t.n = n;
      const int grid = (n + kB - 1) / kB;
      // the e-/e+ part of the stack comes first
      const int ge = (ne + kB - 1) / kB, gg = (n - ne + kB - 1) / kB;
      if (ge > 0) {
        if (fp64) {
          hipLaunchKernelGGL(k_sh_el_substep<double>, dim3(ge), dim3(kB), 0, st, a, g, res, xend, side, HepemTables64(ctx),
                             seed, ne);
          hipLaunchKernelGGL(k_sh_el_end<double>, dim3(ge), dim3(kB), 0, st, a, res, HepemTables64(ctx), seed, ne);
        } else {
          hipLaunchKernelGGL(k_sh_el_substep<float>, dim3(ge), dim3(kB), 0, st, a, g, res, xend, side, HepemTables32(ctx),
                             seed, ne);
          hipLaunchKernelGGL(k_sh_el_end<float>, dim3(ge), dim3(kB), 0, st, a, res, HepemTables32(ctx), seed, ne);
        }
        r->substeps += ne;
      }
      if (gg > 0) {
        if (fp64 && woodcock)
          hipLaunchKernelGGL((k_sh_gamma<double, true>), dim3(gg), dim3(kB), 0, st, a, g, res, xend, side, HepemTables64(ctx), seed, ne);
        else if (fp64)
          hipLaunchKernelGGL((k_sh_gamma<double, false>), dim3(gg), dim3(kB), 0, st, a, g, res, xend, side, HepemTables64(ctx), seed, ne);
        else if (woodcock)
          hipLaunchKernelGGL((k_sh_gamma<float, true>), dim3(gg), dim3(kB), 0, st, a, g, res, xend, side, HepemTables32(ctx), seed, ne);
        else
          hipLaunchKernelGGL((k_sh_gamma<float, false>), dim3(gg), dim3(kB), 0, st, a, g, res, xend, side, HepemTables32(ctx), seed, ne);
      }
      hipLaunchKernelGGL(k_sh_post, dim3(grid), dim3(kB), 0, st, a, g, res, xend, side, dep, leak, keep, smask, offs, offs + cap,
                         bsum, bsum + nblk, bsec);
      // primaries of the next events join once the stack runs low, and no
      // sooner after the last ones than the first events took to reach their
      // largest stack: the stack then stays below about refill plus one chunk's
      // largest stack
      const bool refillNow = refill > 0 && n < refill && next < ns && growRounds >= 0 && sinceInject >= growRounds;
      const int k = refillNow ? ((ns - next) < chunk ? (ns - next) : chunk) : 0;
      ++tag;
      hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanB), 0, st, bsum, bsum + nblk, bsec, grid, k, su->primary != 2 ? 1 : 0,
                         total, nsec, nsec + 1, hc.d, tag);
      hipLaunchKernelGGL(k_sh_scatter, dim3(grid), dim3(kB), 0, st, a, b, g, res, keep, smask, offs, offs + cap, bsum, bsum + nblk,
                         total, cap);
      if (k > 0) {
        hipLaunchKernelGGL(k_sh_inject, dim3((k + kB - 1) / kB), dim3(kB), 0, st, b, g, su->primary, su->ekin0,
                           (int)(s0 + next), next, k, total + 1, total, cap);
        r->tracks += k;
        next += k;
        rounds = 0;
      }
      SH_OK(hipGetLastError());
      // the next round is enqueued as soon as its size is known; stream order
      // keeps it behind this round's scatter
      if (!hc.wait(tag, st)) return HEPEM_ERR_HIP;
      const uint32_t m = hc.count();
      r->iterations += 1;
      if (m > (uint32_t)cap) return HEPEM_ERR_CAPACITY;
      if (growRounds < 0 && (int)m < n) growRounds = sinceInject + 1;  // the stack shrank for the first time
      sinceInject = k > 0 ? 0 : sinceInject + 1;
      n = (int)m;
      ne = (int)hc.electrons();
      cur ^= 1;
    }
    SH_OK(hipMemcpyAsync(hdep.data(), dep, (size_t)ns * nbins * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    SH_OK(hipMemcpyAsync(hleak.data(), leak, (size_t)ns * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    SH_OK(hipStreamSynchronize(st));
    if (r->edep_event)
      for (size_t k = 0; k < (size_t)ns * nbins; ++k) r->edep_event[(size_t)s0 * nbins + k] = (int64_t)hdep[k];
    if (r->leak_event)
      for (int e = 0; e < ns; ++e) r->leak_event[s0 + e] = (int64_t)hleak[e];
    for (int e = 0; e < ns; ++e) {
      for (int k = 0; k < nbins; ++k) {
        const double v = (double)hdep[(size_t)e * nbins + k] / kFix;
        ks[k].add(v);
        ks2[k].add(v * v);
        ktot.add(v);
      }
      kleak.add((double)hleak[e] / kFix);
    }
  }
  SH_OK(hipMemcpyAsync(hnsec, nsec, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
  SH_OK(hipStreamSynchronize(st));
  r->tracks += (int64_t)hnsec[0];
  r->steps = (int64_t)hnsec[1];
  for (int k = 0; k < nbins; ++k) {
    r->edep_sum[k] = ks[k].s;
    r->edep_sum2[k] = ks2[k].s;
  }
  r->leak_sum = kleak.s;
  r->edep_total = ktot.s;
  return HEPEM_OK;
}
