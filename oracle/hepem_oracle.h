/*
 * hepem_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference's run-time stepping physics (G4HepEmRun),
 * used as the correctness oracle for the HIP kernels and as the CPU baseline.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load or call it; the product library never links it.
 */
#ifndef HEPEM_ORACLE_H
#define HEPEM_ORACLE_H

#include <stdint.h>
#include "g4hepem_amd_data.h"
#include "g4hepem_amd.h"
#include "g4hepem_amd_shower.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Philox4x32-10 block (Salmon et al., SC'11). */
void   oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* `n` uniform doubles of the stream (seed, id) starting at draw index `first`. */
void   oracle_rng_uniforms(uint64_t seed, uint32_t id, uint32_t first, int n, double* out);

/* VDT/Cephes exp and log as used on the reference host path. */
double oracle_vdt_exp(double x);
double oracle_vdt_log(double x);

/* config-1 lookups: out has 5*n doubles (dedx, range, invrange(range), xs ioni, xs brem) */
int oracle_lookup_eloss_xsec(const struct G4HepEmData* d, int ptype, int n, const int32_t* imc,
                             const double* ekin, const double* lekin, double* out);

/* Fine-grained probes of single reference functions, n queries each. */
/* which: 0 GetRestMacXSecForStepping ioni, 1 ... brem, 2 GetTransportMFP (imat = couple's material),
          3 GetMacXSecNuclear, 4 ComputeMacXsecAnnihilation, 5 ComputeMacXsecAnnihilationForStepping */
int oracle_probe_electron(const struct G4HepEmData* d, int ptype, int which, int n, const int32_t* imc,
                          const double* ekin, const double* lekin, double* out);
/* element selectors: which 0 brem SB, 1 brem RB (SelectTargetAtom, couple imc), 2 conversion (material of imc) */
int oracle_probe_selector(const struct G4HepEmData* d, int ptype, int which, int n, const int32_t* imc,
                          const double* ekin, const double* lekin, const double* urnd, int32_t* out);
/* gamma: total macroscopic xsec (out_xs), PE xsec stored in the track (out_pe) and
   SampleInteraction winner for urnd (out_pid) */
int oracle_probe_gamma(const struct G4HepEmData* d, int n, const int32_t* imc, const double* ekin,
                       const double* lekin, const double* urnd, double* out_xs, double* out_pe, int32_t* out_pid);

/* One HowFar + Perform step of each track (host arrays, same layout as the
 * device API). out may alias in. Uses up to `nthreads` host threads. */
int oracle_step(const struct G4HepEmData* d, const struct G4HepEmParameters* p, uint64_t seed,
                const hepem_tracks* in, hepem_tracks* out, hepem_step_out* res, int nthreads);
/* The same with a geometry between HowFar and Perform: track i's geometrical
 * step is limited to geom_limit[i] (NULL: no limit); a limited step ends on a
 * boundary (out->on_boundary = 1), any other step does not (0). */
int oracle_step_geom(const struct G4HepEmData* d, const struct G4HepEmParameters* p, uint64_t seed,
                     const hepem_tracks* in, hepem_tracks* out, hepem_step_out* res, const double* geom_limit,
                     int nthreads);

/* One e-/e+ step of each track as G4HepEmTrackingManager::TrackElectron takes it
 * (G4HepEm/G4HepEm/src/G4HepEmTrackingManager.cc:408-635), with MSC sub-steps where
 * the couple's region sets fIsMultipleStepsInMSCTrans, in a slab geometry: track i
 * is at x[i] in the slab lo[i] <= x <= hi[i] across the x axis, and crossing lo[i]
 * (bit 0 of exitf[i]) or hi[i] (bit 1) leaves the world. The pre-step safety is the
 * distance to the nearer plane (0 on a boundary). Updates t in place and x[i] to the
 * post-step position (MSC displacements applied); side[i] = +1/-1 if the step ended
 * on hi/lo, else 0; left[i] = 1 if it left the world; nsub[i] = sub-steps taken.
 * res: true_step = summed true path, geom_step = summed geometrical path,
 * displacement = 0 (applied), the rest as oracle_step. Every track must be an
 * e-/e+ with a valid couple and a positive energy. */
int oracle_electron_slab_step(const struct G4HepEmData* d, const struct G4HepEmParameters* p, uint64_t seed,
                              hepem_tracks* t, hepem_step_out* res, double* x, const double* lo, const double* hi,
                              const uint8_t* exitf, int8_t* side, uint8_t* left, int32_t* nsub, int nthreads);

/* CPU harness of hepem_slab_run: the same tracking loop, per event on up to
 * `nthreads` threads. dep: [nevents][nlayers*nabs] and leak: [nevents] in
 * 1e-9 MeV units; steps_tracks (optional): {track-steps, tracks}. */
int oracle_slab_shower(const struct G4HepEmData* d, const struct G4HepEmParameters* p, const hepem_slab_setup* su,
                       int nevents, uint64_t seed, int64_t* dep, int64_t* leak, int64_t* steps_tracks, int nthreads);

#ifdef __cplusplus
}
#endif

#endif
