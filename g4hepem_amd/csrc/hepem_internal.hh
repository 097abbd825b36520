// hepem_internal.hh -- functions shared between the library's translation
// units that are not part of the C-ABI (hidden in the shared object).
#pragma once

#include "g4hepem_amd.h"

// Production thresholds of couple `imc` in energy -- e-, e+, gamma, as
// G4HepEmMCCData::fSec{El,Pos,Gam}ProdCutE -- and whether the couple's region
// applies them to secondaries (G4HepEmRegionParmeters::fIsApplyCuts of
// region G4HepEmMCCData::fG4RegionIndex). HEPEM_ERR_INVALID for an unknown
// couple.
__attribute__((visibility("hidden"))) int HepemCoupleCuts(const hepem_ctx* ctx, int imc, double cuts[3], int* apply);
