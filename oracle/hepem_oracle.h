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

/* CPU harness of hepem_slab_run: the same tracking loop, per event on up to
 * `nthreads` threads. dep: [nevents][nlayers*nabs] and leak: [nevents] in
 * 1e-9 MeV units; steps_tracks (optional): {track-steps, tracks}. */
int oracle_slab_shower(const struct G4HepEmData* d, const struct G4HepEmParameters* p, const hepem_slab_setup* su,
                       int nevents, uint64_t seed, int64_t* dep, int64_t* leak, int64_t* steps_tracks, int nthreads);

#ifdef __cplusplus
}
#endif

#endif
