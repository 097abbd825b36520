// hepem_validate.hh -- consistency check of a G4HepEmData before anything
// indexes it (not part of the C-ABI; hidden in the shared object).
#pragma once

#include "g4hepem_amd_data.h"

// True when every index the run-time code derives from the tables themselves
// stays inside the arrays that hold it: the material-cuts couples' material
// indices, the elements' Z values (<= fMaxZet), the e-/e+ restricted cross
// section blocks (both the ioni and the brem block behind each start index),
// the element-selector blocks of every multi-element material, the
// Seltzer-Berger rows of every element in use and the gamma-cut indices of
// every couple, the gamma data block sizes and the conversion selector, and
// the per-material counts of the e-/e+ and gamma data. The reference reads
// all of these unchecked (G4HepEm/G4HepEmRun/include/G4HepEmRunUtils.icc:51,
// 224-226); tables read from a document or passed to a context are checked
// here once, so a malformed table set is refused instead of read out of
// bounds on the host or the device.
__attribute__((visibility("hidden"))) bool HepemTablesValid(const G4HepEmData* d);
