/*
 * g4hepem_amd_tables.h -- offline, deterministic builder of G4HepEmData /
 * G4HepEmParameters in the reference layout.
 *
 * The reference fills these tables from Geant4 models in G4HepEmInit
 * (InitG4HepEmState, G4HepEm/G4HepEmInit/src/G4HepEmStateInit.cc:12-28), which
 * needs Geant4 and G4LEDATA. This builder produces the same array shapes,
 * grids and index conventions from analytic models (Berger-Seltzer restricted
 * stopping power, Moller/Bhabha, Tsai complete-screening bremsstrahlung with
 * dielectric suppression, Klein-Nishina, Bethe-Heitler pair production with
 * Coulomb correction, a K/L-edge photoabsorption model, screened-Rutherford
 * transport cross sections, RITA-parametrised bremsstrahlung sampling
 * tables). Table physics is approximate; run-time parity is defined on
 * identical tables.
 */
#ifndef G4HEPEM_AMD_TABLES_H
#define G4HEPEM_AMD_TABLES_H

#include <stdint.h>
#include "g4hepem_amd_data.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Build tables for `nmat` materials (one material-cuts couple each, in the
 * given order, all in detector region 0) with a common production cut given
 * as a range in mm. Known material names: G4_Pb, G4_lAr, G4_Si, G4_PbWO4,
 * G4_W, G4_Cu, G4_Fe, G4_WATER. Returns 0 or a negative hepem_status. */
int hepem_tables_build(const char* const* material_names, int nmat, double prod_cut_mm,
                       struct G4HepEmData** out_data, struct G4HepEmParameters** out_pars);

/* Fill the lepto-nuclear (e-/e+) and photo-nuclear macroscopic cross
 * sections of existing tables. The offline builder leaves them zero, as
 * Geant4 does without hadronic physics; tables in which they are non-zero
 * reach the nuclear branches of the run-time code:
 *   e-/e+: GetMacXSecNuclear and the winner-3 step limit and delta check
 *          (G4HepEm/G4HepEmRun/include/G4HepEmElectronManager.icc:85,418-421,534-541);
 *   gamma: the total cross section of window 2 exceeds the sum of
 *          conversion, Compton and PE, so SampleInteraction can return 3
 *          (G4HepEmGammaManager.icc:199-214).
 * Cross sections are enuc_scale (gnuc_scale) times a fixed smooth energy
 * shape in units of 1/X0 of the material: e-/e+ 1 + ln(E/100 MeV)/4 above
 * 100 MeV; gamma a 13 MeV giant-dipole bump over a plateau above 5 MeV.
 * Spline second derivatives are recomputed. Scales of 0 clear them. */
int hepem_tables_set_nuclear(struct G4HepEmData* data, double enuc_scale, double gnuc_scale);

/* Free everything hepem_tables_build allocated. */
int hepem_tables_free(struct G4HepEmData* data, struct G4HepEmParameters* pars);

/* Serialise every table value (and the parameters) into a flat byte stream in
 * a fixed order. With buf == NULL only *size is set. Used for the digest of
 * the table content. */
int hepem_tables_serialize(const struct G4HepEmData* data, const struct G4HepEmParameters* pars,
                           uint8_t* buf, uint64_t* size);

/* G4HepEmState documents in the reference's JSON schema
 * (G4HepEm/G4HepEmDataJsonIO/src/G4HepEmDataJsonIOImpl.hh; written by
 * G4HepEmStateToJson, G4HepEmDataJsonIO.cc:45-53), so tables that Geant4
 * built elsewhere load unchanged.
 * hepem_tables_from_json replaces G4HepEmStateFromJson (:56-61): it reads a
 * document {"fData": ..., "fParameters": ...} of `len` bytes into newly
 * allocated tables (free them with hepem_tables_free). A document holding
 * only G4HepEmData (G4HepEmDataToJson) and one holding only the parameters
 * (G4HepEmParametersToJson) can be joined into that form as the two values.
 * Returns HEPEM_ERR_INVALID, and allocates nothing, on malformed JSON, a
 * missing member or array sizes that do not fit together.
 * hepem_tables_to_json replaces G4HepEmStateToJson: with buf == NULL only
 * *size is set; otherwise buf holds *size bytes and receives the text (not
 * NUL-terminated), *size its length; HEPEM_ERR_CAPACITY if it does not fit.
 * Doubles are written with 17 significant digits, so reading the text back
 * reproduces every value exactly. */
int hepem_tables_from_json(const char* text, uint64_t len, struct G4HepEmData** out_data,
                           struct G4HepEmParameters** out_pars);
int hepem_tables_to_json(const struct G4HepEmData* data, const struct G4HepEmParameters* pars, char* buf,
                         uint64_t* size);

#ifdef __cplusplus
}
#endif

#endif /* G4HEPEM_AMD_TABLES_H */
