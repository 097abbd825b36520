/*
 * g4hepem_amd_shower.h -- showers in a 1D stack of slabs, stepped on the GPU
 * with the library's HowFar / Perform calls (include/g4hepem_amd.h).
 *
 * No single reference counterpart: this is the tracking loop of
 * G4HepEmTrackingManager::TrackElectron / TrackGamma
 * (G4HepEm/G4HepEm/src/G4HepEmTrackingManager.cc:230-777, 781-1188) for a
 * geometry of planes perpendicular to x, laterally infinite, driven batch-wise
 * over a device track stack instead of one track at a time:
 *   safety -> HowFar -> geometry (cut at the next plane, post-step boundary
 *   flag) -> Perform -> move (pre-step direction, then the MSC displacement
 *   with the reference's safety rule, :527-566) -> score -> stack survivors
 *   and secondaries (secondaries start at the post-step point).
 * Tracks are followed until the physics stops them (tracking cuts) or they
 * leave the stack of slabs. Secondaries are stacked as StackSecondaries does
 * (:1274-1313): where the region of the parent's couple applies cuts
 * (G4HepEmRegionParmeters::fIsApplyCuts), a secondary below that couple's
 * production threshold is deposited at the parent's step instead (an e+ with
 * 2 mc^2 added, when the gamma threshold exceeds the electron mass). There is
 * no sub-stepping of the along-step part (the reference's Geant4 integration
 * loop, :482-606).
 * With woodcock_min_ekin > 0, a gamma above that energy is tracked through the
 * stack as G4HepEmWoodcockHelper::KeepTracking does
 * (G4HepEm/G4HepEm/src/G4HepEmWoodcockHelper.cc:145-300, driven at
 * G4HepEmTrackingManager.cc:955-1062): steps of the densest absorber's mean
 * free path, each point accepted as an interaction with probability
 * Sigma(material there) / Sigma(densest), until an interaction or the stack's
 * exit (less 1e-3 mm) is reached. The interaction then happens in the slab
 * where it lies, with that slab's couple (the densest material's own couple
 * where it lies in that material); a gamma that reaches the exit leaves the
 * stack in that step. A gamma within 1e-3 + 1e-6 mm of the exit is tracked
 * normally. The Woodcock step replaces the gamma's HowFar proposal, so the one
 * number HowFar drew from its stream for the interaction length goes unused.
 *
 * Every track owns the Philox stream selected by a 64-bit id (id, rng_stream)
 * derived from its parent's id, the parent's step number and the secondary
 * slot, so results do not depend on how tracks are batched. Deposits are
 * summed per event and slab in 64-bit fixed point (1e-9 MeV units): exact and
 * independent of the summation order. Per-slab sums over events use
 * compensated (Kahan) FP64 summation.
 */
#ifndef G4HEPEM_AMD_SHOWER_H
#define G4HEPEM_AMD_SHOWER_H

#include <stdint.h>
#include "g4hepem_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

#define HEPEM_SLAB_MAX_ABS 4

typedef struct hepem_slab_setup {
  int32_t nlayers;                      /* repetitions of the absorber sequence           */
  int32_t nabs;                         /* absorbers per repetition (1..4)                 */
  double  thickness[HEPEM_SLAB_MAX_ABS];/* [mm] along x                                    */
  int32_t imc[HEPEM_SLAB_MAX_ABS];      /* material-cuts couple of each absorber           */
  int32_t primary;                      /* hepem_particle of the primaries                 */
  double  ekin0;                        /* primary kinetic energy [MeV], along +x from x=0 */
  int32_t capacity;                     /* track-stack capacity (tracks in flight)         */
  int32_t events_per_chunk;             /* events whose primaries are stacked together     */
  int32_t refill_below;                 /* stack the next events_per_chunk primaries once
                                           fewer tracks than this are in flight, no sooner
                                           after the previous ones than the first events
                                           took to reach their largest stack
                                           (0: only once the stack is empty)               */
  double  woodcock_min_ekin;            /* gammas above this kinetic energy [MeV] are
                                           Woodcock-tracked through the whole stack (the
                                           reference's default limit is 0.2); 0: off       */
} hepem_slab_setup;

typedef struct hepem_slab_result {
  /* caller-provided host arrays of nlayers*nabs entries (slab k = layer*nabs + absorber) */
  double*  edep_sum;                    /* sum over events of the event's deposit [MeV]   */
  double*  edep_sum2;                   /* sum over events of its square [MeV^2]          */
  int64_t* edep_event;                  /* optional [nevents][nlayers*nabs]: each event's
                                           deposits in 1e-9 MeV units (NULL: not kept)     */
  int64_t* leak_event;                  /* optional [nevents], same units                 */
  /* filled by the call */
  double   leak_sum;                    /* energy leaving the stack, summed over events    */
  double   edep_total;                  /* all deposits, summed over events                */
  int64_t  steps;                       /* track-steps                                     */
  int64_t  tracks;                      /* primaries + secondaries                         */
  int32_t  iterations;                  /* stepping rounds over the stack                  */
  int32_t  max_in_flight;               /* largest stack occupancy                         */
  int64_t  substeps;                    /* e-/e+ sub-steps (MSC-limited sub-steps of the
                                           e-/e+ steps, each counted)                      */
} hepem_slab_result;

/* Simulates `nevents` events (event e has primary stream id derived from
 * (seed, e)), stacking their primaries setup->events_per_chunk at a time:
 * into the empty stack, or, with refill_below > 0, beside the tracks still in
 * flight once fewer than refill_below remain (so stepping rounds stay full);
 * the stack then holds at most about refill_below plus the largest stack of one
 * chunk of events. Results do not depend on either setting. Allocates its stack on the
 * context's device for the duration of the call; synchronises `hip_stream`.
 * HEPEM_ERR_CAPACITY if the stack would exceed setup->capacity. */
int hepem_slab_run(const hepem_ctx* ctx, const hepem_slab_setup* setup, int32_t nevents, uint64_t seed,
                   hepem_slab_result* result, void* hip_stream);

#ifdef __cplusplus
}
#endif

#endif /* G4HEPEM_AMD_SHOWER_H */
