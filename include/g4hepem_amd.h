/*
 * g4hepem_amd.h -- C-ABI of the MI355X (gfx950) run-time stepping-physics
 * library. Plain pointers and sizes only; every entry point returns an int
 * status (0 = ok, negative = hepem_status) and never exits the process.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *
 *   hepem_ctx_create / hepem_ctx_destroy
 *       replace CopyG4HepEmDataToGPU / FreeG4HepEmDataOnGPU
 *       (G4HepEm/G4HepEmData/src/G4HepEmData.cc:78-102) and
 *       CopyG4HepEmParametersToGPU (G4HepEm/G4HepEmData/src/G4HepEmParameters.cc:28-36).
 *       The context deep-copies the tables into one device image (FP64 and
 *       optional FP32 copies); it is immutable afterwards.
 *
 *   hepem_lookup_eloss_xsec
 *       the batched form of G4HepEmElectronManager::GetRestDEDX / GetRestRange /
 *       GetInvRange / GetRestMacXSec
 *       (G4HepEm/G4HepEmRun/include/G4HepEmElectronManager.hh:90-150,
 *        .icc:486-532), as driven by the reference test kernels
 *       TestElossDataKernel (testing/ElectronEnergyLoss/src/ELossData.cu:18-26)
 *       and TestResMacXSecDataKernel (testing/ElectronXSections/src/ResMacXSecs.cu:16-23).
 *
 *   hepem_step
 *       one HowFar + Perform step of every track of a batch:
 *       G4HepEmElectronManager::HowFar(data, pars, tlData) + Perform(data, pars, tlData)
 *       (G4HepEmElectronManager.hh:71,222; .icc:35-45,461-483) for e-/e+ and
 *       G4HepEmGammaManager::HowFar + SelectInteraction + Perform
 *       (G4HepEmGammaManager.hh:32-55; .icc:27-35,54-94,173-178) for gammas,
 *       with no geometry between HowFar and Perform (geometric step = physics
 *       step, not on a boundary).
 *
 *   hepem_batch_plan_* / hepem_step_batched
 *       many independent track batches of uneven size in one launch (no
 *       reference counterpart: the reference steps one track per thread,
 *       G4HepEm/G4HepEm/src/G4HepEmTrackingManager.cc:346,895).
 *
 * Random numbers: every track owns a counter-based Philox4x32-10 stream keyed
 * by (seed, track id) -- the 64-bit pair (rng_stream, id) when the optional
 * rng_stream array is given; its draw counter and the cached second Gaussian
 * deviate of G4HepEmRandomEngine::Gauss (G4HepEmRandomEngine.hh:50-70) are
 * part of the track state, so a batch can be re-run or resumed bit-exactly.
 *
 * All track arrays are device pointers (HBM), structure-of-arrays, indexed by
 * track slot i in [0, n). Multi-component fields are stored as planes:
 * component c of track i is at [c*n + i].
 */
#ifndef G4HEPEM_AMD_H
#define G4HEPEM_AMD_H

#include <stdint.h>
#include "g4hepem_amd_data.h"

#ifdef __cplusplus
extern "C" {
#endif

enum hepem_status {
  HEPEM_OK              =  0,
  HEPEM_ERR_INVALID     = -1,  /* null/invalid argument or context */
  HEPEM_ERR_HIP         = -2,  /* a HIP runtime call failed          */
  HEPEM_ERR_CAPACITY    = -3,  /* a size exceeds a fixed capacity     */
  HEPEM_ERR_ALLOC       = -4,  /* device or host allocation failed    */
  HEPEM_ERR_NO_DEVICE   = -5,  /* no usable gfx950 device             */
  HEPEM_ERR_UNSUPPORTED = -6   /* option not available in this build  */
};

enum hepem_precision {
  HEPEM_FP64  = 0,   /* tables and arithmetic in double (bit-exact form)   */
  HEPEM_FP32T = 1    /* FP32 tables/arithmetic with FP64 islands            */
};

enum hepem_particle { HEPEM_ELECTRON = 0, HEPEM_POSITRON = 1, HEPEM_GAMMA = 2 };

typedef struct hepem_ctx hepem_ctx;

/* Per-track state, read and written by the stepping entry points. */
typedef struct hepem_tracks {
  int32_t   n;                 /* number of tracks                              */
  int8_t*   ptype;             /* hepem_particle                                */
  int32_t*  id;                /* track id: selects the Philox stream           */
  int32_t*  imc;               /* material-cuts couple index                    */
  double*   ekin;              /* kinetic energy [MeV]                          */
  double*   dir;               /* [3][n] unit direction                         */
  double*   safety;            /* isotropic safety [mm]                         */
  uint8_t*  on_boundary;       /* pre-step point on a volume boundary           */
  double*   num_ia_left;       /* [4][n] number of interaction lengths left     */
  double*   msc_init_range;    /* Urban MSC persistent state (e-/e+ only)       */
  double*   msc_dyn_range_factor;
  double*   msc_tlimit_min;
  uint8_t*  msc_first_step;
  uint32_t* rng_ctr;           /* Philox draw counter                           */
  double*   rng_gauss;         /* cached second Gaussian deviate                */
  uint8_t*  rng_has_gauss;
  uint32_t* rng_stream;        /* optional (NULL: 0): high word of the 64-bit
                                  stream selector, stream = rng_stream:id   */
} hepem_tracks;

/* Per-step results (write only). Up to 2 secondaries per track and step:
 * slot s < nsec[i] holds one (electrons first, then gammas); slots at or
 * beyond nsec[i] are not written. */
typedef struct hepem_step_out {
  double*   true_step;         /* physical (true) step length [mm]               */
  double*   geom_step;         /* geometrical step length [mm]                   */
  double*   edep;              /* energy deposit in the step [MeV]               */
  double*   displacement;      /* [3][n] MSC lateral displacement (lab frame)     */
  int32_t*  winner;            /* winner process index (reference convention)    */
  uint8_t*  stopped;           /* 1 if the track stopped/was killed in this step  */
  int32_t*  nsec;              /* number of secondaries created (0..2)           */
  int8_t*   sec_ptype;         /* [2][n] hepem_particle of secondary s            */
  double*   sec_ekin;          /* [2][n]                                          */
  double*   sec_dir;           /* [3][2][n]: component c of slot s at (c*2+s)*n+i */
} hepem_step_out;

/* Step-local state that a HowFar call leaves for the Perform call of the same
 * step: the fields of G4HepEmElectronTrack (fRange, fPStepLength:
 * G4HepEmElectronTrack.hh:51-58), G4HepEmTrack (fGStepLength, fMFPs,
 * fWinnerProcessIndex), G4HepEmMSCTrackData (G4HepEmMSCTrackData.hh:112-131)
 * and G4HepEmGammaTrack (fPEmxSec, G4HepEmGammaTrack.hh:32-45) that HowFar
 * computes and Perform reads. */
typedef struct hepem_howfar_state {
  int32_t   n;
  double*   geom_step;         /* HowFar: proposed geometrical step [mm]; the
                                  caller's geometry may shorten it before Perform */
  double*   range;             /* e-/e+: restricted range at the pre-step point   */
  double*   mfp;               /* [4][n] macroscopic mean free paths [mm]
                                  (gamma: [0] = total)                            */
  int32_t*  winner;            /* process that limited the physics step           */
  double*   msc_lambtr1;       /* e-/e+ Urban MSC: first transport mfp            */
  double*   msc_true_step;     /*   true step length                              */
  double*   msc_zpath;         /*   geometrical (z) path length                   */
  double*   msc_par1;          /*   true<->geometrical conversion parameters      */
  double*   msc_par2;
  double*   msc_par3;
  uint8_t*  msc_active;        /*   MSC active in this step                       */
  uint8_t*  msc_displace;      /*   lateral displacement to be sampled            */
  double*   pe_xsec;           /* gamma: photoelectric macroscopic cross section  */
} hepem_howfar_state;

/* What one e-/e+ step carries across its MSC sub-steps when the caller drives
 * G4HepEmTrackingManager::TrackElectron's loop itself
 * (G4HepEm/G4HepEm/src/G4HepEmTrackingManager.cc:427-606): the locals of that
 * loop, per track. Written by the hepem_electron_step_* / _substep_* calls;
 * the caller's geometry sets left_world. */
typedef struct hepem_substep_state {
  int32_t   n;
  int32_t*  discrete_winner;   /* the process HowFarToDiscreteInteraction chose (iDProc) */
  double*   step_left;         /* what is left of the discrete step (stepLimitLeft) */
  double*   true_path;         /* summed true sub-step lengths (totalTruePathLength) */
  double*   geom_path;         /* summed geometrical sub-step lengths              */
  double*   mean_eloss;        /* summed mean energy losses (totalEloss)           */
  double*   pre_ekin;          /* kinetic energy at the step's pre-step point      */
  double*   displacement;      /* [3][n] MSC displacement of the last sub-step (lab
                                  frame; 0 when none was sampled)                  */
  uint8_t*  active;            /* 1: the track takes the current sub-step          */
  uint8_t*  more;              /* 1: another sub-step follows (continueStepping)   */
  uint8_t*  stopped;           /* the track stopped in a sub-step                  */
  uint8_t*  left_world;        /* set by the caller: the last move left the world  */
  int32_t*  nsub;              /* sub-steps taken                                  */
} hepem_substep_state;

/* One batch of the batched launch. */
typedef struct hepem_batch_desc {
  hepem_tracks   in;
  hepem_tracks   out;          /* may alias `in` (in-place update)               */
  hepem_step_out res;
} hepem_batch_desc;

/* ---- context ------------------------------------------------------------ */
int  hepem_ctx_create(const struct G4HepEmData* data, const struct G4HepEmParameters* pars,
                      int precision, hepem_ctx** out_ctx);
int  hepem_ctx_destroy(hepem_ctx* ctx);
const char* hepem_status_string(int status);
/* Device bytes held by the context's table image. */
int  hepem_ctx_table_bytes(const hepem_ctx* ctx, uint64_t* bytes);


/* ---- config-1 style lookups ---------------------------------------------
 * For n queries (imc, E, lnE) of particle `ptype` (e- or e+) writes
 * out[0*n+i] restricted dE/dx, out[1*n+i] range, out[2*n+i] inverse range of
 * that range, out[3*n+i] ioni and out[4*n+i] brem restricted macroscopic
 * cross sections. `out` holds 5*n doubles. All pointers are device pointers. */
int  hepem_lookup_eloss_xsec(const hepem_ctx* ctx, int ptype, int32_t n,
                             const int32_t* imc, const double* ekin, const double* lekin,
                             double* out, void* hip_stream);

/* ---- gamma cross-section lookups -------------------------------------------
 * For n gammas (imc, E, lnE) and uniform draws u: out_xs[i] the total
 * macroscopic cross section (G4HepEmGammaManager::GetTotalMacXSec,
 * G4HepEm/G4HepEmRun/include/G4HepEmGammaManager.icc:108-153), out_pid[i] the
 * interaction SampleInteraction (.icc:187-219) selects for u with the mean
 * free path 1/out_xs[i], and out_pe[i] the photoelectric cross section the
 * track holds afterwards -- the batched form of the reference test kernel
 * TestMacXSecDataKernel (testing/GammaXSections/src/GammaMacXSecs.cu:20-31).
 * All pointers are device pointers. */
int  hepem_lookup_gamma_xsec(const hepem_ctx* ctx, int32_t n, const int32_t* imc, const double* ekin,
                             const double* lekin, const double* urnd, double* out_xs, double* out_pe,
                             int32_t* out_pid, void* hip_stream);

/* ---- stepping ------------------------------------------------------------
 * One step of every track of the batch. `out` may alias `in`. `work` must be
 * at least hepem_step_work_bytes(n) bytes of device memory (step counters,
 * the particle-type queue, a process code and a staged record per slot, and
 * the tile-local process layout of the pending interactions);
 * pass NULL to let the context use its own buffer (then calls on one context
 * must not overlap). */
int  hepem_step(const hepem_ctx* ctx, uint64_t seed, const hepem_tracks* in,
                hepem_tracks* out, hepem_step_out* res, void* work, void* hip_stream);
uint64_t hepem_step_work_bytes(int32_t n);

/* hepem_step with no work buffer and no allocation of any kind. No counterpart
 * in the reference: this entry point is for callers that cannot allocate
 * device memory -- code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc. The step's counters, particle-type queue
 * and staged interactions live in a fixed array that the library holds in
 * each GPU's memory from load until the process ends
 * (hepem_step_work_bytes(HEPEM_NOALLOC_MAX_TRACKS) bytes, about 17.7 MiB);
 * its device address is looked up when a context is created on that GPU.
 * Batches above HEPEM_NOALLOC_MAX_TRACKS tracks return HEPEM_ERR_CAPACITY.
 * Every call on one GPU uses that GPU's one array: calls that overlap on a
 * GPU (other streams, threads or contexts) race and must be serialised by the
 * caller. Each call clears the counters it needs with a kernel (no memset),
 * so what an earlier call left in the array does not matter. Results are
 * bit-identical to hepem_step on the same inputs. The number of rejected
 * tracks is not reported (hepem_step_rejected needs a work buffer). */
#define HEPEM_NOALLOC_MAX_TRACKS 262144
int  hepem_step_noalloc(const hepem_ctx* ctx, uint64_t seed, const hepem_tracks* in, hepem_tracks* out,
                        hepem_step_out* res, void* hip_stream);

/* The two phases of hepem_step, for callers that time them separately:
 * hepem_step_prepare clears the step counters and lays the track slots out
 * by particle type in the queue held in `work`: with the process sort, within
 * each tile of 4096 slots (its e-/e+ first, then its gammas); without it,
 * over the whole batch (e-/e+ first). The layout is recorded in `work`, so
 * hepem_step_run steps the queue correctly even if the process-sort mode
 * changed between the two calls; hepem_step_run
 * performs the physics step of every queued slot: HowFar, the continuous step
 * and the choice of the discrete interaction, then (see
 * hepem_ctx_set_process_sort) a second sort of the pending interactions by
 * process and the final-state sampling with whole waves per process. `work` is required here. hepem_step == prepare +
 * run on the same stream. */
int  hepem_step_prepare(const hepem_ctx* ctx, const hepem_tracks* in, void* work, void* hip_stream);
int  hepem_step_run(const hepem_ctx* ctx, uint64_t seed, const hepem_tracks* in, hepem_tracks* out,
                    hepem_step_out* res, void* work, void* hip_stream);
/* hepem_step_run, batch plans and the split Perform calls (hepem_perform,
 * hepem_*_perform) in one fused kernel (HEPEM_SORT_OFF), with the process sort
 * (HEPEM_SORT_ON), or with the faster of the two for the call as measured on
 * MI355X (HEPEM_SORT_AUTO, the default), which is the process sort for every
 * call at every size in both forms. Results are identical; only the speed
 * differs. */
#define HEPEM_SORT_OFF  0
#define HEPEM_SORT_ON   1
#define HEPEM_SORT_AUTO 2
int  hepem_ctx_set_process_sort(hepem_ctx* ctx, int mode);
/* A speed switch of hepem_lookup_eloss_xsec, like the process sort above:
 * the particle's loss and cross-section tables staged in LDS (on = 1, the
 * default; used where they take at most 64 kB) or read from the table image
 * (on = 0). Results are identical. (The lookups themselves follow
 * G4HepEmElectronManager's GetRestRange, GetRestDEDX, GetInvRange and
 * GetRestMacXSec, G4HepEm/G4HepEmRun/include/G4HepEmElectronManager.icc:486-532.) */
int  hepem_ctx_set_lookup_staging(hepem_ctx* ctx, int on);

/* Number of tracks the last step on `work` rejected (invalid couple index or
 * non-positive/non-finite energy; such tracks get winner -99 and are left
 * unchanged). Synchronises `hip_stream`. */
int  hepem_step_rejected(const void* work, uint32_t* nrejected, void* hip_stream);

/* ---- HowFar and Perform as separate calls, for callers with geometry -----
 * hepem_howfar: G4HepEmElectronManager::HowFar (G4HepEmElectronManager.hh:71;
 * .icc:35-45) for the e-/e+ and G4HepEmGammaManager::HowFar
 * (G4HepEmGammaManager.hh:32; .icc:27-48) for the gammas of a mixed batch.
 * It reads the pre-step `on_boundary` and `safety`, updates the persistent
 * state of `t` in place (interaction lengths, MSC state, RNG) and fills `hs`.
 * The caller then moves the tracks: it may shorten hs->geom_step and sets
 * t->on_boundary to the post-step status.
 * hepem_perform: G4HepEmElectronManager::Perform (.hh:222; .icc:461-483) and,
 * for gammas, the boundary update or SelectInteraction + Perform of
 * G4HepEmGammaManager (.hh:38-55; .icc:54-105,173-219), as driven by
 * G4HepEmTrackingManager.cc:999-1000,1092-1108. Updates `t` in place and fills
 * `res`. With hs->geom_step and on_boundary left unchanged, howfar + perform
 * equals hepem_step bit for bit.
 * `work` (hepem_step_work_bytes(n) bytes, or NULL for the context's own
 * buffer) holds the particle-type queue built by hepem_howfar; pass the same
 * buffer to the hepem_perform of the same step. Tracks that hepem_howfar
 * rejects get hs->winner = -99 and hs->geom_step = 0. */
int  hepem_howfar(const hepem_ctx* ctx, uint64_t seed, hepem_tracks* t, hepem_howfar_state* hs,
                  void* work, void* hip_stream);
int  hepem_perform(const hepem_ctx* ctx, uint64_t seed, hepem_tracks* t, const hepem_howfar_state* hs,
                   hepem_step_out* res, void* work, void* hip_stream);
/* The same for batches of one kind: every slot must hold an e-/e+
 * (electron calls) or a gamma (gamma calls); other slots are rejected. No
 * particle queue is built; `work` holds the rejection counter and, with the
 * process sort, the staged interactions (NULL allowed). */
int  hepem_electron_howfar(const hepem_ctx* ctx, uint64_t seed, hepem_tracks* t, hepem_howfar_state* hs,
                           void* work, void* hip_stream);
int  hepem_electron_perform(const hepem_ctx* ctx, uint64_t seed, hepem_tracks* t, const hepem_howfar_state* hs,
                            hepem_step_out* res, void* work, void* hip_stream);
int  hepem_gamma_howfar(const hepem_ctx* ctx, uint64_t seed, hepem_tracks* t, hepem_howfar_state* hs,
                        void* work, void* hip_stream);
int  hepem_gamma_perform(const hepem_ctx* ctx, uint64_t seed, hepem_tracks* t, const hepem_howfar_state* hs,
                         hepem_step_out* res, void* work, void* hip_stream);

/* ---- TrackElectron's MSC sub-steps, for callers with their own geometry ----
 * G4HepEmTrackingManager::TrackElectron (G4HepEm/G4HepEm/src/G4HepEmTrackingManager.cc:
 * 408-635) does not call HowFar and Perform: it drives the fine-grained
 * G4HepEmElectronManager functions (G4HepEm/G4HepEmRun/include/G4HepEmElectronManager.hh:
 * 90-206) and, where the region sets fIsMultipleStepsInMSCTrans, repeats the
 * MSC part of the step in sub-steps until the discrete step is used up. A
 * caller with that loop binds to these calls, on batches of e-/e+ (other slots
 * are rejected: hs->winner = -99, ss->active = 0):
 *
 *   hepem_electron_step_begin       pre-step point: the interaction lengths
 *                                   and HowFarToDiscreteInteraction (:427-443);
 *                                   t->safety must be the pre-step safety (0
 *                                   on a boundary) and stays so for the step;
 *   loop while hepem_electron_substep_active reports tracks:
 *     hepem_electron_substep_howfar HowFarToMSC of the active tracks
 *                                   (:446-452): hs->geom_step = proposed move;
 *     (caller)                      moves each active track at most
 *                                   hs->geom_step, writes the length moved to
 *                                   hs->geom_step and the post-move boundary
 *                                   flag to t->on_boundary, sets
 *                                   ss->left_world for a track that left the
 *                                   world (:456-486);
 *     hepem_electron_substep_along  UpdatePStepLength, UpdateNumIALeft,
 *                                   ApplyMeanEnergyLoss, SampleMSC (:490-521)
 *                                   and the next sub-step's set-up (:569-594);
 *                                   ss->displacement is the MSC displacement
 *                                   the caller applies off a boundary
 *                                   (:523-566);
 *   hepem_electron_step_end         the loss fluctuation on the summed mean
 *                                   loss (:597-606), at-rest annihilation or
 *                                   PerformDiscrete (:611-650, winner 3: the
 *                                   interaction-length reset and the delta test,
 *                                   the caller's nuclear final state); fills res
 *                                   (true_step = summed true path, geom_step =
 *                                   summed moves, displacement = 0).
 *
 * `work`: device memory of at least hepem_step_work_bytes(0) bytes for the
 * loop's counters (NULL: the context's own buffer, whose count
 * hepem_electron_substep_active cannot read).
 * With sub-steps off in every region and no track leaving the world, begin +
 * substep_howfar + substep_along + end equal hepem_electron_howfar +
 * hepem_electron_perform with the same geometry bit for bit. */
int  hepem_electron_step_begin(const hepem_ctx* ctx, uint64_t seed, hepem_tracks* t, hepem_howfar_state* hs,
                               hepem_substep_state* ss, void* work, void* hip_stream);
int  hepem_electron_substep_howfar(const hepem_ctx* ctx, uint64_t seed, hepem_tracks* t, hepem_howfar_state* hs,
                                   hepem_substep_state* ss, void* work, void* hip_stream);
int  hepem_electron_substep_along(const hepem_ctx* ctx, uint64_t seed, hepem_tracks* t, hepem_howfar_state* hs,
                                  hepem_substep_state* ss, void* work, void* hip_stream);
/* Tracks that take another sub-step after the last hepem_electron_step_begin
 * or hepem_electron_substep_along on `work`. Synchronises the stream. */
int  hepem_electron_substep_active(const void* work, uint32_t* nactive, void* hip_stream);
int  hepem_electron_step_end(const hepem_ctx* ctx, uint64_t seed, hepem_tracks* t, const hepem_howfar_state* hs,
                             hepem_substep_state* ss, hepem_step_out* res, void* work, void* hip_stream);

/* ---- many independent batches in one launch -------------------------------
 * A batch plan holds, in device memory, the descriptors of nbatch batches of
 * uneven size, their particle-type queues and per-batch counters. Creating it
 * allocates and uploads (synchronously); running it only enqueues kernels
 * (counter fill, one batched classify, one batched step) -- no allocation, no
 * host copies -- so a run can be captured into a hipGraph and replayed.
 * Results are bit-identical to calling hepem_step on each batch with the same
 * seed. A plan sorts pending interactions by process (as hepem_step_run) as
 * its context's mode at plan creation says for the plan's total slot count
 * (hepem_ctx_set_process_sort). The batches' arrays must stay valid while the
 * plan is used. */
typedef struct hepem_batch_plan hepem_batch_plan;
int  hepem_batch_plan_create(const hepem_ctx* ctx, const hepem_batch_desc* descs, int32_t nbatch,
                             hepem_batch_plan** out_plan);
int  hepem_batch_plan_destroy(hepem_batch_plan* plan);
int  hepem_batch_plan_run(const hepem_batch_plan* plan, uint64_t seed, void* hip_stream);
/* Launch shape of a plan's step. Every shape gives the same results.
 *   block       threads per block of the step kernel: 64, 128 or 256;
 *   sort_tile   slots per tile of the process sort of the pending
 *               interactions (1024, 2048, 4096 or 8192), or 0: no process
 *               sort (one fused kernel);
 *   persistent  0: one block per chunk of work, dispatched by the hardware;
 *               1: one resident round of blocks taking chunks from a work
 *               counter (reset by every run) until none are left;
 *   grid        (read only) blocks of the step kernel, or of its first stage
 *               with the process sort: one per chunk of `block` slots, or,
 *               dispatched at 256 threads, 16 per 4096-slot tile of the
 *               batches (tiles rounded up to a multiple of 8), or the
 *               persistent form's resident round. */
typedef struct hepem_batch_shape {
  int32_t block;
  int32_t sort_tile;
  int32_t persistent;
  int32_t grid;
} hepem_batch_shape;
/* Times every launch shape the plan may take on the plan's own batches and
 * keeps the fastest: each block size; with and without the process sort as
 * the context's mode at plan creation allows (HEPEM_SORT_AUTO: both); each
 * sort tile size; dispatched and persistent. This runs the step several
 * times per shape: batches stepped in place advance. Synchronises the
 * stream. */
int  hepem_batch_plan_tune(hepem_batch_plan* plan, uint64_t seed, void* hip_stream);
/* Launch shape in use: block size and grid blocks of the step kernel. */
int  hepem_batch_plan_shape(const hepem_batch_plan* plan, int32_t* block, int32_t* grid);
/* The full launch shape in use, and setting one (grid is ignored on input;
 * HEPEM_ERR_INVALID for a block, tile or persistent value not listed above). */
int  hepem_batch_plan_get_shape(const hepem_batch_plan* plan, hepem_batch_shape* shape);
int  hepem_batch_plan_set_shape(hepem_batch_plan* plan, const hepem_batch_shape* shape);
/* Tracks of batch `b` the last run rejected (see hepem_step_rejected).
 * Synchronises the stream. */
int  hepem_batch_plan_rejected(const hepem_batch_plan* plan, int32_t b, uint32_t* nrejected, void* hip_stream);

/* Convenience form: create a plan for `descs` (a host array), run it once,
 * wait for it and free it. Allocates; not for graph capture. */
int  hepem_step_batched(const hepem_ctx* ctx, uint64_t seed, const hepem_batch_desc* descs,
                        int32_t nbatch, void* hip_stream);

/* ---- a device track stack -------------------------------------------------
 * For callers that step batches with their own geometry. The current batch
 * lives in the stack's device memory: hepem_stack_tracks gives the view to
 * step, in place or into another SoA. hepem_stack_advance then builds the
 * next batch from that step's output, as G4HepEmTrackingManager::
 * StackSecondaries does for one track
 * (G4HepEm/G4HepEm/src/G4HepEmTrackingManager.cc:1254-1330):
 *   - a track stays if keep[i] != 0 (keep NULL: every track; the caller's
 *     geometry clears it for a track that left the world) and its kinetic
 *     energy is above zero;
 *   - each secondary of its step (res: nsec, sec_*) is stacked, unless the
 *     region of the parent's couple (out->imc[i]) applies production cuts and
 *     the secondary is an e- below the e- threshold, an e+ below the e+
 *     threshold while the gamma threshold exceeds the electron mass, or a
 *     gamma below the gamma threshold; such a secondary is deposited instead
 *     (an e+ with its annihilation energy 2 m_e c^2), into cut_edep[i]
 *     (optional, out->n entries);
 *   - the next batch holds the e-/e+ first, then the gammas, each in parent
 *     order: a track before its stacked secondaries, those in slot order.
 *     parent[j] and origin[j] (optional, capacity entries each) give the slot
 *     of the step's output that slot j came from and whether it is that track
 *     (0) or its secondary slot 0 or 1 (1, 2), so the caller can carry its own
 *     per-track data (position, touchable) along;
 *   - a stacked secondary starts at its parent's post-step point (the
 *     parent's couple, safety and boundary flag) with the start-of-track
 *     state of hepem_tracks_reset_state and a Philox stream of its own,
 *     derived from the parent's stream, its post-step draw counter and the
 *     slot.
 * `out` may be the view of the current batch but not the stack's other
 * buffer. Advancing synchronises the stream to read the next count
 * (*n_next); HEPEM_ERR_CAPACITY (next count above the capacity) leaves the
 * current batch as it was. hepem_stack_push appends a caller's tracks to the
 * current batch (device arrays, copied; rng_stream may be NULL). */
typedef struct hepem_stack hepem_stack;
int  hepem_stack_create(const hepem_ctx* ctx, int32_t capacity, hepem_stack** out_stack);
int  hepem_stack_destroy(hepem_stack* stack);
int  hepem_stack_tracks(const hepem_stack* stack, hepem_tracks* view);
int  hepem_stack_push(hepem_stack* stack, const hepem_tracks* tracks, void* hip_stream);
/* Copies the current batch into the caller's SoA (dst->n = the batch's count). */
int  hepem_stack_copy_out(const hepem_stack* stack, hepem_tracks* dst, void* hip_stream);
int  hepem_stack_advance(hepem_stack* stack, const hepem_tracks* out, const hepem_step_out* res, const uint8_t* keep,
                         double* cut_edep, int32_t* parent, int8_t* origin, int32_t* n_next, void* hip_stream);

/* ---- track-state helpers -------------------------------------------------
 * Reset the persistent physics state of n tracks to the reference's start-of-
 * track values (G4HepEmTrack::ReSet, G4HepEmMSCTrackData::ReSet,
 * G4HepEmRandomEngine::DiscardGauss): num_ia_left = -1, MSC first-step
 * defaults, rng_ctr = 0, no cached Gaussian. */
int  hepem_tracks_reset_state(const hepem_tracks* t, void* hip_stream);

#ifdef __cplusplus
}
#endif

#endif /* G4HEPEM_AMD_H */
