// hepem_internal.hh -- functions shared between the library's translation
// units that are not part of the C-ABI (hidden in the shared object).
#pragma once

#include "device/dev_common.hh"
#include "g4hepem_amd.h"

// Production thresholds of couple `imc` in energy -- e-, e+, gamma, as
// G4HepEmMCCData::fSec{El,Pos,Gam}ProdCutE -- and whether the couple's region
// applies them to secondaries (G4HepEmRegionParmeters::fIsApplyCuts of
// region G4HepEmMCCData::fG4RegionIndex). HEPEM_ERR_INVALID for an unknown
// couple.
__attribute__((visibility("hidden"))) int HepemCoupleCuts(const hepem_ctx* ctx, int imc, double cuts[3], int* apply);

// The context's precision (HEPEM_FP64 or HEPEM_FP32T) and the two forms of its
// device tables (valid while the context lives).
__attribute__((visibility("hidden"))) int HepemPrecision(const hepem_ctx* ctx);
__attribute__((visibility("hidden"))) const hepem::DevTables<double>& HepemTables64(const hepem_ctx* ctx);
__attribute__((visibility("hidden"))) const hepem::DevTables<float>& HepemTables32(const hepem_ctx* ctx);

// G4HepEm material index of couple `imc` and that material's density
// (G4HepEmMatData::fDensity). HEPEM_ERR_INVALID for an unknown couple.
__attribute__((visibility("hidden"))) int HepemCoupleMaterial(const hepem_ctx* ctx, int imc, int* imat, double* density);
